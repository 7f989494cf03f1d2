// Host side of the frame-element group: geometry of LinearCrdTransf2d/3d,
// fiber-section definitions and the Steel02 / Concrete02 / Elastic state
// pools, uploaded once at setup; the batched operations launch the kernels
// of xg_frame.hip.
#include "xg_frame.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>

#include "xg_frame_dev.h"

namespace xg {

// Gauss-Lobatto locations on [-1, 1] and weights for 2..10 points as
// LobattoBeamIntegration::getSectionLocations / getSectionWeights tabulate
// them (LobattoBeamIntegration.cpp:113-293); the elements use 0.5 (x + 1)
// and 0.5 w.
static bool lobatto_points(int n, double* xi, double* wt) {
  static const double X[11][10] = {
      {}, {},
      {-1.0, 1.0},
      {-1.0, 0.0, 1.0},
      {-1.0, -0.44721360, 0.44721360, 1.0},
      {-1.0, -0.65465367, 0.0, 0.65465367, 1.0},
      {-1.0, -0.7650553239, -0.2852315164, 0.2852315164, 0.7650553239, 1.0},
      {-1.0, -0.8302238962, -0.4688487934, 0.0, 0.4688487934, 0.8302238962, 1.0},
      {-1.0, -0.8717401485, -0.5917001814, -0.2092992179, 0.2092992179, 0.5917001814, 0.8717401485, 1.0},
      {-1.0, -0.8997579954, -0.6771862795, -0.3631174638, 0.0, 0.3631174638, 0.6771862795, 0.8997579954, 1.0},
      {-1.0, -0.9195339082, -0.7387738651, -0.4779249498, -0.1652789577, 0.1652789577, 0.4779249498,
       0.7387738651, 0.9195339082, 1.0}};
  static const double W[11][10] = {
      {}, {},
      {1.0, 1.0},
      {0.333333333333333, 1.333333333333333, 0.333333333333333},
      {0.166666666666667, 0.833333333333333, 0.833333333333333, 0.166666666666667},
      {0.1, 0.5444444444, 0.7111111111, 0.5444444444, 0.1},
      {0.06666666667, 0.3784749562, 0.5548583770, 0.5548583770, 0.3784749562, 0.06666666667},
      {0.04761904762, 0.2768260473, 0.4317453812, 0.4876190476, 0.4317453812, 0.2768260473, 0.04761904762},
      {0.03571428571, 0.2107042271, 0.3411226924, 0.4124587946, 0.4124587946, 0.3411226924, 0.2107042271,
       0.03571428571},
      {0.02777777778, 0.1654953615, 0.2745387125, 0.3464285109, 0.3715192743, 0.3464285109, 0.2745387125,
       0.1654953615, 0.02777777778},
      {0.02222222222, 0.1333059908, 0.2248893421, 0.2920426836, 0.3275397611, 0.3275397611, 0.2920426836,
       0.2248893421, 0.1333059908, 0.02222222222}};
  if (n < 2 || n > 10) return false;
  for (int i = 0; i < n; i++) {
    xi[i] = 0.5 * (X[n][i] + 1.0);
    wt[i] = W[n][i] * 0.5;
  }
  return true;
}

// host Gauss-Jordan inverse with partial pivoting (the device's inverse_n)
static bool host_inverse(int n, const double* A, double* Ai) {
  double a[6][12];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) a[i][j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int i = c + 1; i < n; i++)
      if (std::fabs(a[i][c]) > std::fabs(a[piv][c])) piv = i;
    if (a[piv][c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) std::swap(a[c][j], a[piv][j]);
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * n; j++) a[c][j] *= inv;
    for (int i = 0; i < n; i++)
      if (i != c) {
        const double fct = a[i][c];
        for (int j = 0; j < 2 * n; j++) a[i][j] -= fct * a[c][j];
      }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Ai[i * n + j] = a[i][n + j];
  return true;
}

struct FrameGroup : Group {
  bool d3 = false;
  bool force = false;  // ForceBeamColumn2d/3d
  int nip = 5, maxiter = 10;
  double tol = 1e-12;
  std::vector<int> sec_h, transf_h;
  std::vector<double> rho_h;
  FrameDev f;
  DVec<double> dlxi, dlwt, dfs0, dSe, dSe_c, dkv, dkv_c, dvprev, dvprev_c, dvs, dvs_c, dfsec, dSsr;
  DVec<int> dinit, dferr;
  int NB() const { return d3 ? 6 : 3; }
  // device storage
  DVec<int> dconn, dsecdef, ddef_off, dfkind, dfmat, dflocal, drun_off, drun_beg, drun_kind, drun_mat, ddef_last;
  DVec<double> dgeo, dfy, dfz, dfA, dGJ, de, ds, dks, delasE;
  DVec<long long> dsb, dcb, deb;
  DVec<int> drsec;
  DVec<Steel02Param> dsteelp;
  DVec<Concrete02Param> dconcp;
  DVec<double> st_c, st_t, cs_c, cs_t, el_c, el_t;
  DVec<int> kon_c, kon_t;
  std::vector<double> geo_h;
  int ord() const { return d3 ? 4 : 2; }
  int ld() const { return d3 ? 4 : 2; }
  int npts() const override { return nip; }
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
    const long long nsec = (long long)ne * nip;
    std::vector<long long> sb(nsec), cb(nsec), eb(nsec);
    long long ns = 0, nc = 0, nl = 0;
    nfib = 0;
    for (int e = 0; e < ne; e++)
      for (int i = 0; i < nip; i++) {
        const long long sec = (long long)e * nip + i;
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
        for (int i = 0; i < nip; i++) {
          const long long sec = (long long)e * nip + i;
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
    f.nip = nip;
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
    // the displacement-based update in rows (section_rows): passes per
    // wavefront of each definition for rows of 16 and 20 lanes -- a run of
    // n fibers takes ceil(n / rw) passes over the row (a Steel02 pass counted
    // twice, for its two powers; an elastic one a quarter), the section
    // deformation and totals one.  XG_FRAME_ROWS=16 or 20 fixes the width,
    // =0 keeps one wavefront per section.
    const char* rows_env = std::getenv("XG_FRAME_ROWS");
    rows_on = !force && nsec > 0 && !(rows_env && rows_env[0] == '0');
    if (rows_on) {
      const int fixed = rows_env ? std::atoi(rows_env) : 0;
      rows_fixed = (fixed == 16 || fixed == 20) ? fixed : 0;
      rows_secdef = secdef;
      rows_pass16.assign(GJ.size(), 1.0);
      rows_pass20.assign(GJ.size(), 1.0);
      for (size_t df = 0; df < GJ.size(); df++)
        for (int r = run_off[df]; r < run_off[df + 1]; r++) {
          const int nf = run_beg[r + 1] - run_beg[r];
          const double wk = run_kind[r] == 1 ? 2.0 : run_kind[r] == 2 ? 1.0 : 0.25;
          rows_pass16[df] += wk * ((nf + 15) / 16);
          rows_pass20[df] += wk * ((nf + 19) / 20);
        }
      build_rows({0, ne}, s);
    }
    if (force) setup_force(d, def_off, fkind, fmat, fy, fz, fA, GJ, steelp, concp, elasE);
    stc_init = std::move(stc);
    csc_init = std::move(csc);
    konc_init = std::move(konc);
    // Domain::addElement: setDomain + update() at zero displacement
    // (DispBeamColumn2d.cpp:184; ForceBeamColumn2d's first update forms the
    // initial basic stiffness)
    zero_update(d);
  }

  // Element and section state of the force-based elements, the Lobatto rule
  // and each definition's initial section flexibility (the inverse of the
  // fibers' initial tangents, SectionForceDeformation::getInitialFlexibility).
  void setup_force(Domain& d, const std::vector<int>& def_off, const std::vector<int>& fkind,
                   const std::vector<int>& fmat, const std::vector<double>& fy, const std::vector<double>& fz,
                   const std::vector<double>& fA, const std::vector<double>& GJ,
                   const std::vector<Steel02Param>& steelp, const std::vector<Concrete02Param>& concp,
                   const std::vector<double>& elasE) {
    hipStream_t s = d.stream;
    const int ne = nele(), nb = NB(), ns = ord();
    const long long nsec = (long long)ne * nip;
    std::vector<double> xi(nip), wt(nip);
    if (!lobatto_points(nip, xi.data(), wt.data()))
      throw Error(XG_ERR_ARG, "forceBeamColumn: 2 to 10 Lobatto points");
    const int ndef = (int)GJ.size();
    std::vector<double> fs0((size_t)ndef * ns * ns, 0.0);
    for (int df = 0; df < ndef; df++) {
      double k[16] = {0};
      for (int fi = def_off[df]; fi < def_off[df + 1]; fi++) {
        const double E = fkind[fi] == 1 ? steelp[fmat[fi]].E0
                                         : (fkind[fi] == 2 ? 2.0 * concp[fmat[fi]].fc / concp[fmat[fi]].epsc0
                                                           : elasE[fmat[fi]]);
        const double y = fy[fi], A = fA[fi];
        if (!d3) {
          const double ks0 = E * A, ks1 = ks0 * -y;
          k[0] += ks0;
          k[1] += ks1;
          k[3] += ks1 * -y;
        } else {
          const double z = fz[fi], EA = E * A;
          k[0] += EA;
          k[1] += -y * EA;
          k[2] += z * EA;
          k[5] += y * y * EA;
          k[6] += -y * z * EA;
          k[10] += z * z * EA;
        }
      }
      if (!d3) {
        k[2] = k[1];
      } else {
        if (GJ[df] < 0.0) throw Error(XG_ERR_ARG, "forceBeamColumn (3D): the section needs a torsional stiffness (-GJ)");
        k[4] = k[1];
        k[8] = k[2];
        k[9] = k[6];
        k[15] = GJ[df];
      }
      double kk[16];
      for (int a = 0; a < ns; a++)
        for (int b = 0; b < ns; b++) kk[a * ns + b] = k[a * (d3 ? 4 : 2) + b];
      if (!host_inverse(ns, kk, &fs0[(size_t)df * ns * ns]))
        throw Error(XG_ERR_ARG, "forceBeamColumn: singular initial section stiffness");
    }
    dlxi.upload(xi, s);
    dlwt.upload(wt, s);
    dfs0.upload(fs0, s);
    auto zeros = [&](DVec<double>& v, size_t n) {
      v.resize(n);
      v.zero(s);
    };
    zeros(dSe, (size_t)ne * nb);
    zeros(dSe_c, (size_t)ne * nb);
    zeros(dkv, (size_t)ne * nb * nb);
    zeros(dkv_c, (size_t)ne * nb * nb);
    zeros(dvprev, (size_t)ne * nb);
    zeros(dvprev_c, (size_t)ne * nb);
    zeros(dvs, (size_t)nsec * ns);
    zeros(dvs_c, (size_t)nsec * ns);
    zeros(dfsec, (size_t)nsec * ns * ns);
    zeros(dSsr, (size_t)nsec * ns);
    dinit.resize(ne);
    dinit.zero(s);
    dferr.resize(1);
    dferr.zero(s);
    f.force = 1;
    f.maxiter = maxiter;
    f.tol = tol;
    f.lxi = dlxi.p;
    f.lwt = dlwt.p;
    f.fs0 = dfs0.p;
    f.Se = dSe.p;
    f.Se_c = dSe_c.p;
    f.kv = dkv.p;
    f.kv_c = dkv_c.p;
    f.vprev = dvprev.p;
    f.vprev_c = dvprev_c.p;
    f.init = dinit.p;
    f.vs = dvs.p;
    f.vs_c = dvs_c.p;
    f.fsec = dfsec.p;
    f.Ssr = dSsr.p;
  }
  void force_zero_state(hipStream_t s) {
    for (DVec<double>* v : {&dSe, &dSe_c, &dkv, &dkv_c, &dvprev, &dvprev_c, &dvs, &dvs_c, &dfsec, &dSsr}) v->zero(s);
    dinit.zero(s);
  }
  std::vector<double> stc_init, csc_init;
  std::vector<int> konc_init;

  // rows of the displacement-based update (section_rows)
  bool rows_on = false;
  int rows_fixed = 0;                            // XG_FRAME_ROWS=16 / 20, else 0
  std::vector<int> rows_secdef;                  // section definition of each element
  std::vector<double> rows_pass16, rows_pass20;  // wavefront passes per definition

  // Sections grouped by (part, definition), the parts being the element
  // ranges [parts[p], parts[p+1]), 64 / rw of a group to a wavefront, the
  // row width rw the one of 16 and 20 with fewer passes in all (rows left
  // empty at the end of a group counted); returns the first wavefront of
  // each part and the end (empty if rows are off).
  std::vector<int> build_rows(const std::vector<int>& parts, hipStream_t s) {
    std::vector<int> off;
    if (!rows_on) return off;
    std::vector<std::pair<int, std::vector<int>>> groups;  // (definition, sections)
    std::vector<size_t> part_first;
    for (size_t p = 0; p + 1 < parts.size(); p++) {
      part_first.push_back(groups.size());
      std::map<int, size_t> at;
      for (int e = parts[p]; e < parts[p + 1]; e++) {
        const int df = rows_secdef[e];
        auto it = at.find(df);
        if (it == at.end()) {
          it = at.emplace(df, groups.size()).first;
          groups.push_back({df, {}});
        }
        for (int i = 0; i < nip; i++) groups[it->second].second.push_back(e * nip + i);
      }
    }
    part_first.push_back(groups.size());
    auto cost = [&](int w) {
      const size_t nr = 64 / w;
      double c = 0.0;
      for (const auto& g : groups) c += (double)((g.second.size() + nr - 1) / nr) * (w == 20 ? rows_pass20 : rows_pass16)[g.first];
      return c;
    };
    const int rw = rows_fixed ? rows_fixed : (cost(20) < cost(16) ? 20 : 16);
    const size_t nr = 64 / rw;
    std::vector<int> rsec;
    for (size_t p = 0; p + 1 < part_first.size(); p++) {
      off.push_back((int)(rsec.size() / nr));
      for (size_t g = part_first[p]; g < part_first[p + 1]; g++) {
        const std::vector<int>& v = groups[g].second;
        for (size_t i = 0; i < v.size(); i += nr)
          for (size_t k = 0; k < nr; k++) rsec.push_back(i + k < v.size() ? v[i + k] : -1);
      }
    }
    off.push_back((int)(rsec.size() / nr));
    if (rsec.empty()) rsec.push_back(-1);
    drsec.upload(rsec, s);
    f.rsec = drsec.p;
    f.nrw = off.back();
    f.rw = rw;
    return off;
  }

  void zero_update(Domain& d) {
    DVec<double> zeroU;
    zeroU.resize((size_t)d.nnode * d.ndf);
    zeroU.zero(d.stream);
    if (force) launch_force_update(d.stream, f, d3, zeroU.p, dferr.p);
    else launch_frame_update(d.stream, f, d3, zeroU.p, d.nfail.p);
    XG_HIP(hipStreamSynchronize(d.stream));
  }
  void update(Domain& d) override {
    if (force) launch_force_update(d.stream, f, d3, d.U.p, dferr.p);
    else launch_frame_update(d.stream, f, d3, d.U.p, d.nfail.p);
  }
  bool update_failed(Domain& d) override {
    if (!force) return false;
    int h = 0;
    dferr.download(&h, 1, d.stream);
    if (h) dferr.zero(d.stream);
    return h != 0;
  }
  void resid(Domain& d, double* Re) override {
    if (force) launch_force_resid(d.stream, f, d3, Re);
    else launch_frame_resid(d.stream, f, d3, Re);
  }
  void tangent(Domain& d, double* Ke) override {
    if (force) launch_force_tangent(d.stream, f, d3, Ke);
    else launch_frame_tangent(d.stream, f, d3, Ke);
  }
  void commit(Domain& d) override {
    launch_copy_d(d.stream, (long long)st_t.n, st_t.p, st_c.p);
    launch_copy_i(d.stream, (long long)kon_t.n, kon_t.p, kon_c.p);
    launch_copy_d(d.stream, (long long)cs_t.n, cs_t.p, cs_c.p);
    launch_copy_d(d.stream, (long long)el_t.n, el_t.p, el_c.p);
    if (force) {  // ForceBeamColumn2d::commitState (:276-308)
      launch_copy_d(d.stream, (long long)dvs.n, dvs.p, dvs_c.p);
      launch_copy_d(d.stream, (long long)dkv.n, dkv.p, dkv_c.p);
      launch_copy_d(d.stream, (long long)dSe.n, dSe.p, dSe_c.p);
      launch_copy_d(d.stream, (long long)dvprev.n, dvprev.p, dvprev_c.p);
    }
  }
  void revert(Domain& d) override {
    if (force) launch_force_revert(d.stream, f, d3);
    else launch_frame_revert(d.stream, f, d3);
  }
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
    if (force) force_zero_state(s);  // ForceBeamColumn2d::revertToStart (:344-376)
    else zero_update(d);
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
    if (force && (kind == XG_STATE_STEEL02 || kind == XG_STATE_CONCRETE02))
      throw Error(XG_ERR_ARG, "committed-state export: forceBeamColumn element state is not exported");
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

std::vector<int> frame_rows_by_parts(Group& g, const std::vector<int>& ele_off, hipStream_t s) {
  auto* fg = dynamic_cast<FrameGroup*>(&g);
  if (!fg) throw Error(XG_ERR_ARG, "not a frame-element group");
  return fg->build_rows(ele_off, s);
}

std::unique_ptr<Group> make_frame_group(Domain& d, bool force) {
  auto g = std::make_unique<FrameGroup>();
  g->d3 = d.ndm == 3;
  g->force = force;
  g->nen = 2;
  g->nd = 2 * d.ndf;
  return g;
}

void frame_group_add(Domain& d, Group& g, int tag, const EleDef& ed) {
  (void)d;
  (void)tag;
  FrameGroup& fg = static_cast<FrameGroup&>(g);
  if (fg.force) {
    if (fg.tags.size() == 1) {  // the group's first element sets its rule and iteration limits
      fg.nip = ed.nip;
      fg.maxiter = ed.maxiter;
      fg.tol = ed.tol;
    } else if (ed.nip != fg.nip || ed.maxiter != fg.maxiter || ed.tol != fg.tol)
      throw Error(XG_ERR_ARG, "forceBeamColumn: all elements must share nip, maxIters and tol");
  }
  fg.sec_h.push_back(ed.sec);
  fg.transf_h.push_back(ed.transf);
  fg.rho_h.push_back(ed.rho);
}

}  // namespace xg
