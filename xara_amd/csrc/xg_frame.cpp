// Host side of the frame-element group: geometry of LinearCrdTransf2d/3d,
// fiber-section definitions and the Steel02 / Concrete02 / Elastic state
// pools, uploaded once at setup; the batched operations launch the kernels
// of xg_frame.hip.
#include "xg_frame.h"

#include <algorithm>
#include <cmath>

#include "xg_frame_dev.h"

namespace xg {

struct FrameGroup : Group {
  bool d3 = false;
  std::vector<int> sec_h, transf_h;
  std::vector<double> rho_h;
  FrameDev f;
  // device storage
  DVec<int> dconn, dsecdef, ddef_off, dfkind, dfmat, dflocal, drun_off, drun_beg, drun_kind, drun_mat, ddef_last;
  DVec<double> dgeo, dfy, dfz, dfA, dGJ, de, ds, dks, delasE;
  DVec<long long> dsb, dcb, deb;
  DVec<Steel02Param> dsteelp;
  DVec<Concrete02Param> dconcp;
  DVec<double> st_c, st_t, cs_c, cs_t, el_c, el_t;
  DVec<int> kon_c, kon_t;
  std::vector<double> geo_h;
  int ord() const { return d3 ? 4 : 2; }
  int ld() const { return d3 ? 4 : 2; }
  int npts() const override { return 5; }
  long long nfib = 0;
  long long fibers() const override { return nfib; }

  void setup(Domain& d) override {
    hipStream_t s = d.stream;
    const int ne = nele();
    // ---- geometry (computeElemtLengthAndOrient / getLocalAxes) ----
    geo_h.assign((size_t)ne * 10, 0.0);
    for (int e = 0; e < ne; e++) {
      const double* ci = &d.crd_h[(size_t)conn[2 * e] * d.ndm];
      const double* cj = &d.crd_h[(size_t)conn[2 * e + 1] * d.ndm];
      double* g = &geo_h[(size_t)e * 10];
      if (!d3) {
        const double dx0 = cj[0] - ci[0], dx1 = cj[1] - ci[1];
        const double L = std::sqrt(dx0 * dx0 + dx1 * dx1);
        if (L == 0.0) throw Error(XG_ERR_ARG, "dispBeamColumn " + std::to_string(tags[e]) + ": zero length");
        g[0] = L;
        g[1] = dx0 / L;
        g[2] = dx1 / L;
      } else {
        const double dx[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
        const double L = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        if (L == 0.0) throw Error(XG_ERR_ARG, "dispBeamColumn " + std::to_string(tags[e]) + ": zero length");
        const std::vector<double>& v = d.transfdefs.at(transf_h[e]);
        const double x[3] = {dx[0] / L, dx[1] / L, dx[2] / L};
        double y[3] = {v[1] * x[2] - v[2] * x[1], v[2] * x[0] - v[0] * x[2], v[0] * x[1] - v[1] * x[0]};
        const double yn = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        if (yn == 0) throw Error(XG_ERR_ARG, "geomTransf: vecxz is parallel to the element axis");
        for (double& c : y) c /= yn;
        const double z[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
        g[0] = L;
        for (int c = 0; c < 3; c++) {
          g[1 + c] = x[c];
          g[4 + c] = y[c];
          g[7 + c] = z[c];
        }
      }
    }
    // ---- section definitions ----
    std::map<int, int> def_of;
    std::vector<int> def_off{0}, fkind, fmat, flocal, secdef(ne);
    std::vector<int> run_off{0}, run_beg, run_kind, run_mat, def_last;
    std::vector<double> fy, fz, fA, GJ;
    std::vector<int> nsteel_def, nconc_def, nelas_def;
    std::map<int, int> steel_idx, conc_idx, elas_idx;
    std::vector<Steel02Param> steelp;
    std::vector<Concrete02Param> concp;
    std::vector<double> elasE;
    for (auto& kv : d.steeldefs) {
      steel_idx[kv.first] = (int)steelp.size();
      const auto& p = kv.second;
      steelp.push_back({p.Fy, p.E0, p.b, p.R0, p.cR1, p.cR2, p.a1, p.a2, p.a3, p.a4, p.sigini});
    }
    for (auto& kv : d.concdefs) {
      conc_idx[kv.first] = (int)concp.size();
      const auto& p = kv.second;
      concp.push_back({p.fc, p.epsc0, p.fcu, p.epscu, p.rat, p.ft, p.Ets});
    }
    for (auto& kv : d.elasticdefs) {
      elas_idx[kv.first] = (int)elasE.size();
      elasE.push_back(kv.second);
    }
    for (int e = 0; e < ne; e++) {
      auto it = def_of.find(sec_h[e]);
      if (it == def_of.end()) {
        const FiberSecDef& sd = d.secdefs.at(sec_h[e]);
        // centroid as FiberSection2d/3d::addFiber recomputes it (:151-156, :338-345)
        double ABar = 0, QzBar = 0, QyBar = 0, yBar = 0, zBar = 0;
        if (sd.centroid)
          for (size_t i = 0; i < sd.A.size(); i++) {
            ABar += sd.A[i];
            QzBar += sd.y[i] * sd.A[i];
            QyBar += sd.z[i] * sd.A[i];
            yBar = QzBar / ABar;
            zBar = QyBar / ABar;
          }
        // pool slot of every fiber in definition order
        const int nfd = (int)sd.A.size();
        int ns = 0, nc = 0, nl = 0;
        std::vector<int> mat_of(nfd), local_of(nfd);
        for (int i = 0; i < nfd; i++) {
          if (sd.kind[i] == 1) {
            mat_of[i] = steel_idx.at(sd.mat[i]);
            local_of[i] = ns++;
          } else if (sd.kind[i] == 2) {
            mat_of[i] = conc_idx.at(sd.mat[i]);
            local_of[i] = nc++;
          } else {
            mat_of[i] = elas_idx.at(sd.mat[i]);
            local_of[i] = nl++;
          }
        }
        // device order: runs of one (kind, material), definition order within
        // a run, so a wavefront pass updates a single material
        std::vector<int> order(nfd);
        for (int i = 0; i < nfd; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
          return sd.kind[a] != sd.kind[b] ? sd.kind[a] < sd.kind[b] : mat_of[a] < mat_of[b];
        });
        const int base = (int)fA.size();
        for (int k = 0; k < nfd; k++) {
          const int i = order[k];
          if (k == 0 || sd.kind[i] != sd.kind[order[k - 1]] || mat_of[i] != mat_of[order[k - 1]]) {
            run_beg.push_back(base + k);
            run_kind.push_back(sd.kind[i]);
            run_mat.push_back(mat_of[i]);
          }
          if (i == nfd - 1) def_last.push_back(k);
          fy.push_back(sd.y[i] - yBar);
          fz.push_back(d3 ? sd.z[i] - zBar : 0.0);
          fA.push_back(sd.A[i]);
          fkind.push_back(sd.kind[i]);
          fmat.push_back(mat_of[i]);
          flocal.push_back(local_of[i]);
        }
        if (nfd == 0) def_last.push_back(-1);
        run_off.push_back((int)run_beg.size());
        nsteel_def.push_back(ns);
        nconc_def.push_back(nc);
        nelas_def.push_back(nl);
        GJ.push_back(d3 ? sd.GJ : -1.0);
        def_off.push_back((int)fA.size());
        it = def_of.emplace(sec_h[e], (int)GJ.size() - 1).first;
      }
      secdef[e] = it->second;
    }
    // ---- pools ----
    const long long nsec = (long long)ne * 5;
    std::vector<long long> sb(nsec), cb(nsec), eb(nsec);
    long long ns = 0, nc = 0, nl = 0;
    nfib = 0;
    for (int e = 0; e < ne; e++)
      for (int i = 0; i < 5; i++) {
        const long long sec = (long long)e * 5 + i;
        const int df = secdef[e];
        sb[sec] = ns;
        cb[sec] = nc;
        eb[sec] = nl;
        ns += nsteel_def[df];
        nc += nconc_def[df];
        nl += nelas_def[df];
        nfib += def_off[df + 1] - def_off[df];
      }
    // revertToStart states (Steel02.cpp:69-95, Concrete02.cpp:321-337)
    std::vector<double> stc((size_t)kSteelHist * ns), csc((size_t)kConcHist * nc);
    std::vector<int> konc(ns, 0);
    {
      // steel fibers in pool order
      for (int e = 0; e < ne; e++)
        for (int i = 0; i < 5; i++) {
          const long long sec = (long long)e * 5 + i;
          const int df = secdef[e];
          for (int fi = def_off[df]; fi < def_off[df + 1]; fi++) {
            if (fkind[fi] == 1) {
              const Steel02Param& p = steelp[fmat[fi]];
              const long long k = sb[sec] + flocal[fi];
              const double epsy = p.Fy / p.E0;
              double hist[kSteelHist] = {-epsy, epsy, 0, 0, 0, 0, 0, 0, 0, p.E0};
              if (p.sigini != 0.0) {
                hist[7] = p.sigini / p.E0;
                hist[8] = p.sigini;
              }
              for (int c = 0; c < kSteelHist; c++) stc[(size_t)c * ns + k] = hist[c];
            } else if (fkind[fi] == 2) {
              const Concrete02Param& p = concp[fmat[fi]];
              const long long k = cb[sec] + flocal[fi];
              double hist[kConcHist] = {0, 0, 0, 0, 2.0 * p.fc / p.epsc0};
              for (int c = 0; c < kConcHist; c++) csc[(size_t)c * nc + k] = hist[c];
            }
          }
        }
    }
    dconn.upload(conn, s);
    dgeo.upload(geo_h, s);
    dsecdef.upload(secdef, s);
    ddef_off.upload(def_off, s);
    dfy.upload(fy, s);
    dfz.upload(fz, s);
    dfA.upload(fA, s);
    dfkind.upload(fkind, s);
    dfmat.upload(fmat, s);
    dflocal.upload(flocal, s);
    run_beg.push_back((int)fA.size());
    drun_off.upload(run_off, s);
    drun_beg.upload(run_beg, s);
    drun_kind.upload(run_kind, s);
    drun_mat.upload(run_mat, s);
    ddef_last.upload(def_last, s);
    dGJ.upload(GJ, s);
    dsb.upload(sb, s);
    dcb.upload(cb, s);
    deb.upload(eb, s);
    dsteelp.upload(steelp, s);
    dconcp.upload(concp, s);
    delasE.upload(elasE, s);
    st_c.upload(stc, s);
    st_t.upload(stc, s);
    kon_c.upload(konc, s);
    kon_t.upload(konc, s);
    cs_c.upload(csc, s);
    cs_t.upload(csc, s);
    el_c.resize(std::max(nl, 1LL));
    el_c.zero(s);
    el_t.resize(std::max(nl, 1LL));
    el_t.zero(s);
    de.resize((size_t)ord() * nsec);
    de.zero(s);
    ds.resize((size_t)ord() * nsec);
    ds.zero(s);
    dks.resize((size_t)ld() * ld() * nsec);
    dks.zero(s);
    f = FrameDev{};
    f.nele = ne;
    f.nip = 5;
    f.nsec = nsec;
    f.conn = dconn.p;
    f.geo = dgeo.p;
    f.secdef = dsecdef.p;
    f.def_off = ddef_off.p;
    f.fy = dfy.p;
    f.fz = dfz.p;
    f.fA = dfA.p;
    f.fkind = dfkind.p;
    f.fmat = dfmat.p;
    f.flocal = dflocal.p;
    f.run_off = drun_off.p;
    f.run_beg = drun_beg.p;
    f.run_kind = drun_kind.p;
    f.run_mat = drun_mat.p;
    f.def_last = ddef_last.p;
    f.GJ = dGJ.p;
    f.steel_base = dsb.p;
    f.conc_base = dcb.p;
    f.elas_base = deb.p;
    f.e = de.p;
    f.s = ds.p;
    f.ks = dks.p;
    f.nsteel = ns;
    f.nconc = nc;
    f.nelas = nl;
    f.steelp = dsteelp.p;
    f.concp = dconcp.p;
    f.elasE = delasE.p;
    f.st_c = st_c.p;
    f.st_t = st_t.p;
    f.kon_c = kon_c.p;
    f.kon_t = kon_t.p;
    f.cs_c = cs_c.p;
    f.cs_t = cs_t.p;
    f.el_c = el_c.p;
    f.el_t = el_t.p;
    stc_init = std::move(stc);
    csc_init = std::move(csc);
    konc_init = std::move(konc);
    // setDomain calls update() (DispBeamColumn2d.cpp:184): zero displacement
    zero_update(d);
  }
  std::vector<double> stc_init, csc_init;
  std::vector<int> konc_init;

  void zero_update(Domain& d) {
    DVec<double> zeroU;
    zeroU.resize((size_t)d.nnode * d.ndf);
    zeroU.zero(d.stream);
    launch_frame_update(d.stream, f, d3, zeroU.p, d.nfail.p);
    XG_HIP(hipStreamSynchronize(d.stream));
  }
  void update(Domain& d) override { launch_frame_update(d.stream, f, d3, d.U.p, d.nfail.p); }
  void resid(Domain& d, double* Re) override { launch_frame_resid(d.stream, f, d3, Re); }
  void tangent(Domain& d, double* Ke) override { launch_frame_tangent(d.stream, f, d3, Ke); }
  void commit(Domain& d) override {
    launch_copy_d(d.stream, (long long)st_t.n, st_t.p, st_c.p);
    launch_copy_i(d.stream, (long long)kon_t.n, kon_t.p, kon_c.p);
    launch_copy_d(d.stream, (long long)cs_t.n, cs_t.p, cs_c.p);
    launch_copy_d(d.stream, (long long)el_t.n, el_t.p, el_c.p);
  }
  void revert(Domain& d) override { launch_frame_revert(d.stream, f, d3); }
  void revert_to_start(Domain& d) override {
    hipStream_t s = d.stream;
    st_c.upload(stc_init, s);
    st_t.upload(stc_init, s);
    kon_c.upload(konc_init, s);
    kon_t.upload(konc_init, s);
    cs_c.upload(csc_init, s);
    cs_t.upload(csc_init, s);
    el_c.zero(s);
    el_t.zero(s);
    zero_update(d);
  }
  void mass_diag(Domain& d, std::vector<double>& out) override {
    // lumped translational mass 0.5 rho L (DispBeamColumn2d::getMass, cMass = 0)
    out.assign((size_t)nele() * nd, 0.0);
    const int per = d3 ? 6 : 3, ntr = d3 ? 3 : 2;
    for (int e = 0; e < nele(); e++) {
      const double m = 0.5 * rho_h[e] * geo_h[(size_t)e * 10];
      for (int a = 0; a < 2; a++)
        for (int j = 0; j < ntr; j++) out[(size_t)e * nd + a * per + j] = m;
    }
  }
  // Steel02::sendSelf (Steel02.cpp:287-308) state part: epsminP, epsmaxP,
  // epsplP, epss0P, sigs0P, epssrP, sigsrP, konP, epsP, sigP, eP;
  // Concrete02::sendSelf (Concrete02.cpp:339-354): ecminP, deptP, epsP,
  // sigP, eP.  Fibers by element (ascending tag), section, then definition
  // order: the pool order.
  long long state_size(int kind) const override {
    if (kind == XG_STATE_STEEL02) return f.nsteel * 11;
    if (kind == XG_STATE_CONCRETE02) return f.nconc * 5;
    return 0;
  }
  void get_state(Domain& d, int kind, double* out) override {
    if (kind == XG_STATE_STEEL02) {
      const long long n = f.nsteel;
      const std::vector<double> h = st_c.download(d.stream);
      const std::vector<int> k = kon_c.download(d.stream);
      for (long long i = 0; i < n; i++) {
        double* o = out + i * 11;
        for (int c = 0; c < 7; c++) o[c] = h[(size_t)c * n + i];
        o[7] = k[i];
        for (int c = 7; c < 10; c++) o[c + 1] = h[(size_t)c * n + i];
      }
    } else if (kind == XG_STATE_CONCRETE02) {
      const long long n = f.nconc;
      const std::vector<double> h = cs_c.download(d.stream);
      for (long long i = 0; i < n; i++)
        for (int c = 0; c < 5; c++) out[i * 5 + c] = h[(size_t)c * n + i];
    }
  }
  void set_state(Domain& d, int kind, const double* in) override {
    if (kind == XG_STATE_STEEL02) {
      const long long n = f.nsteel;
      std::vector<double> h((size_t)kSteelHist * n);
      std::vector<int> k(n);
      for (long long i = 0; i < n; i++) {
        const double* o = in + i * 11;
        for (int c = 0; c < 7; c++) h[(size_t)c * n + i] = o[c];
        k[i] = (int)o[7];
        for (int c = 7; c < 10; c++) h[(size_t)c * n + i] = o[c + 1];
      }
      st_c.upload(h, d.stream);
      st_t.upload(h, d.stream);
      kon_c.upload(k, d.stream);
      kon_t.upload(k, d.stream);
    } else if (kind == XG_STATE_CONCRETE02) {
      const long long n = f.nconc;
      std::vector<double> h((size_t)kConcHist * n);
      for (long long i = 0; i < n; i++)
        for (int c = 0; c < 5; c++) h[(size_t)c * n + i] = in[i * 5 + c];
      cs_c.upload(h, d.stream);
      cs_t.upload(h, d.stream);
    }
    XG_HIP(hipStreamSynchronize(d.stream));
  }

  int point_data(Domain& d, int what, std::vector<double>& out) override {
    const long long nsec = f.nsec;
    const int o = ord();
    if (what == XG_OUT_STRESS || what == XG_OUT_STRAIN) {
      std::vector<double> h = (what == XG_OUT_STRESS ? ds : de).download(d.stream);
      out.resize((size_t)nsec * o);
      for (long long sc = 0; sc < nsec; sc++)
        for (int c = 0; c < o; c++) out[sc * o + c] = h[c * nsec + sc];
      return o;
    }
    if (what == XG_OUT_TANGENT) {
      std::vector<double> h = dks.download(d.stream);
      out.resize((size_t)nsec * o * o);
      for (long long sc = 0; sc < nsec; sc++)
        for (int a = 0; a < o; a++)
          for (int b = 0; b < o; b++) out[(sc * o + a) * o + b] = h[(long long)(b * ld() + a) * nsec + sc];
      return o * o;
    }
    throw Error(XG_ERR_ARG, "unsupported output for frame elements");
  }
};

FrameDev frame_dev(Group& g) {
  auto* fg = dynamic_cast<FrameGroup*>(&g);
  if (!fg) throw Error(XG_ERR_ARG, "not a frame-element group");
  return fg->f;
}

std::unique_ptr<Group> make_frame_group(Domain& d) {
  auto g = std::make_unique<FrameGroup>();
  g->d3 = d.ndm == 3;
  g->nen = 2;
  g->nd = 2 * d.ndf;
  return g;
}

void frame_group_add(Domain& d, Group& g, int tag, const EleDef& ed) {
  (void)d;
  (void)tag;
  FrameGroup& fg = static_cast<FrameGroup&>(g);
  fg.sec_h.push_back(ed.sec);
  fg.transf_h.push_back(ed.transf);
  fg.rho_h.push_back(ed.rho);
}

}  // namespace xg
