// Host driver of the GPU state-determination path: model definition, setup
// (numbering, CSR pattern, gather-assembly maps), the batched element
// operations, the PCG LinearSOE and the Newton / LoadControl /
// DisplacementControl analysis loop over device-resident vectors.
#include "xg_domain.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "xg_frame.h"
#include "xg_j2.h"
#include "xg_kernels.h"
#include "xg_sym_layout.h"

namespace xg {

// ---------------------------------------------------------------------------
// J2 continuum groups: stdBrick (3D) and FourNodeQuad (plane strain)
// ---------------------------------------------------------------------------
struct J2Group : Group {
  bool brick = true;
  std::vector<int> mat_h;        // index into the J2 table
  std::vector<double> b_h;       // body force per element
  std::vector<double> thk_h;     // quad thickness
  bool has_b = false;
  DVec<int> dconn, dmat;
  DVec<J2Param> dprm;
  DVec<double> db, dthk;
  DVec<double> epn, xin, ept, xit, sig;
  // brick: the tangent coefficients c2, c3 of every point (the tangent is
  // re-formed from them and the stress, xg_j2.h); quad: the upper triangle
  // of the plane-strain tangent
  DVec<double> D;
  int ncomp() const { return brick ? 6 : 4; }
  int nsig() const { return brick ? 6 : 3; }
  int ntan() const { return brick ? 2 : 6; }
  int npts() const override { return brick ? 8 : 4; }
  long long np() const { return (long long)nele() * npts(); }

  void setup(Domain& d) override {
    hipStream_t s = d.stream;
    dconn.upload(conn, s);
    dmat.upload(mat_h, s);
    std::vector<J2Param> tab;
    for (auto& kv : d.j2defs) tab.push_back(kv.second);
    dprm.upload(tab, s);
    has_b = std::any_of(b_h.begin(), b_h.end(), [](double v) { return v != 0.0; });
    if (has_b) db.upload(b_h, s);
    if (!brick) dthk.upload(thk_h, s);
    const long long n = np();
    epn.resize(ncomp() * n);
    ept.resize(ncomp() * n);
    xin.resize(n);
    xit.resize(n);
    sig.resize(nsig() * n);
    D.resize(ntan() * n);
    initial_state(d);
  }
  void update(Domain& d) override {
    if (brick)
      launch_brick_update(d.stream, nele(), dconn.p, dmat.p, dprm.p, d.crd.p, d.U.p, d.ndf, d.dt, np(),
                          epn.p, xin.p, ept.p, xit.p, sig.p, D.p, d.nfail.p);
    else
      launch_quad_update(d.stream, nele(), dconn.p, dmat.p, dprm.p, d.crd.p, d.U.p, d.ndf, d.dt, np(),
                         epn.p, xin.p, ept.p, xit.p, sig.p, D.p, d.nfail.p);
  }
  void resid(Domain& d, double* Re) override {
    if (brick) launch_brick_resid(d.stream, nele(), dconn.p, d.crd.p, has_b ? db.p : nullptr, np(), sig.p, Re);
    else launch_quad_resid(d.stream, nele(), dconn.p, d.crd.p, dthk.p, has_b ? db.p : nullptr, np(), sig.p, Re);
  }
  void tangent(Domain& d, double* Ke) override {
    if (brick) launch_brick_tangent(d.stream, nele(), dconn.p, dmat.p, dprm.p, d.crd.p, np(), sig.p, D.p, Ke);
    else launch_quad_tangent(d.stream, nele(), dconn.p, d.crd.p, dthk.p, np(), D.p, Ke);
  }
  bool tangent_csr(Domain& d) override {
    if (!brick || !d.fz_ok) return false;
    launch_brick_tangent_csr(d.stream, d.num.neq / 3, d.rowptr.p, d.fz_ptr.p, d.fz_ele.p, d.fz_a.p, d.fz_tgt.p,
                             dconn.p, dmat.p, dprm.p, d.crd.p, np(), sig.p, D.p, d.vals.p);
    return true;
  }
  void commit(Domain& d) override { launch_j2_commit(d.stream, np(), ncomp(), ept.p, xit.p, epn.p, xin.p); }
  void revert(Domain&) override {}  // J2Plasticity::revertToLastCommit is a no-op (J2Plasticity.cpp:524)
  void revert_to_start(Domain& d) override {
    // J2Plasticity::revertToStart -> zero (J2Plasticity.cpp:213-224, :532-543):
    // plastic state, stress and strain; the tangent is formed again by the
    // update at zero displacement that ends Domain::revertToStart.
    epn.zero(d.stream);
    ept.zero(d.stream);
    xin.zero(d.stream);
    xit.zero(d.stream);
    sig.zero(d.stream);
    XG_HIP(hipStreamSynchronize(d.stream));
  }
  void initial_state(Domain& d) {
    // J2Plasticity::zero + the constructor's plastic_integrator at zero strain
    // (J2Plasticity.cpp:78-80): zero state and stress, elastic tangent
    // (filled by an update at U = 0).
    epn.zero(d.stream);
    ept.zero(d.stream);
    xin.zero(d.stream);
    xit.zero(d.stream);
    sig.zero(d.stream);
    DVec<double> zeroU;
    zeroU.resize((size_t)d.nnode * d.ndf);
    zeroU.zero(d.stream);
    if (brick)
      launch_brick_update(d.stream, nele(), dconn.p, dmat.p, dprm.p, d.crd.p, zeroU.p, d.ndf, 0.0, np(),
                          epn.p, xin.p, ept.p, xit.p, sig.p, D.p, d.nfail.p);
    else
      launch_quad_update(d.stream, nele(), dconn.p, dmat.p, dprm.p, d.crd.p, zeroU.p, d.ndf, 0.0, np(),
                         epn.p, xin.p, ept.p, xit.p, sig.p, D.p, d.nfail.p);
    XG_HIP(hipStreamSynchronize(d.stream));
  }
  int point_data(Domain& d, int what, std::vector<double>& out) override {
    const long long n = np();
    const int P = npts();
    auto soa = [&](DVec<double>& a, int nc) {
      std::vector<double> h = a.download(d.stream);
      return std::make_pair(h, nc);
    };
    if (what == XG_OUT_STRESS) {
      auto [h, nc] = soa(sig, nsig());
      out.resize((size_t)n * nc);
      for (long long p = 0; p < n; p++)
        for (int c = 0; c < nc; c++) out[p * nc + c] = h[c * n + p];
      return nc;
    }
    if (what == XG_OUT_TANGENT) {
      auto [h, nt] = soa(D, ntan());
      const int m = brick ? 6 : 3;
      out.resize((size_t)n * m * m);
      std::vector<double> st, tab_d;
      std::vector<J2Param> tab;
      if (brick) {
        st.resize((size_t)n * 6);
        XG_HIP(hipMemcpyAsync(st.data(), sig.p, st.size() * sizeof(double), hipMemcpyDeviceToHost, d.stream));
        XG_HIP(hipStreamSynchronize(d.stream));
        for (auto& kv : d.j2defs) tab.push_back(kv.second);
      }
      for (long long p = 0; p < n; p++) {
        double Dp[21];
        if (brick) {
          double s6[6];
          for (int c = 0; c < 6; c++) s6[c] = st[c * n + p];
          j2_tangent_from_stress(tab[mat_h[p / 8]], s6, h[p], h[n + p], Dp);
        }
        for (int a = 0; a < m; a++)
          for (int b = a; b < m; b++) {
            const double v = brick ? Dp[tri_index(m, a, b)] : h[(long long)tri_index(m, a, b) * n + p];
            out[p * m * m + a * m + b] = v;
            out[p * m * m + b * m + a] = v;
          }
      }
      (void)nt;
      return m * m;
    }
    if (what == XG_OUT_PLASTIC) {
      std::vector<double> e = ept.download(d.stream), x = xit.download(d.stream);
      out.assign((size_t)n * 7, 0.0);
      for (long long p = 0; p < n; p++) {
        if (brick)
          for (int c = 0; c < 6; c++) out[p * 7 + c] = e[c * n + p];
        else
          for (int c = 0; c < 4; c++) out[p * 7 + c] = e[c * n + p];
        out[p * 7 + 6] = x[p];
      }
      return 7;
    }
    (void)P;
    throw Error(XG_ERR_ARG, "unsupported output for J2 elements");
  }
  size_t device_bytes() const override {
    return (epn.n + ept.n + xin.n + xit.n + sig.n + D.n) * 8 + dconn.n * 4;
  }
  // J2Plasticity::sendSelf state part (J2Plasticity.cpp:597-614): xi_n, then
  // epsilon_p_n(i, j) row by row; points in element order
  long long state_size(int kind) const override { return kind == XG_STATE_J2 ? np() * 10 : 0; }
  void get_state(Domain& d, int kind, double* out) override {
    if (kind != XG_STATE_J2) return;
    const long long n = np();
    std::vector<double> e = epn.download(d.stream), x = xin.download(d.stream);
    for (long long p = 0; p < n; p++) {
      double t[6] = {0, 0, 0, 0, 0, 0};  // 00 11 22 01 12 20
      for (int c = 0; c < ncomp(); c++) t[c] = e[c * n + p];
      const double m[9] = {t[0], t[3], t[5], t[3], t[1], t[4], t[5], t[4], t[2]};
      out[p * 10] = x[p];
      for (int k = 0; k < 9; k++) out[p * 10 + 1 + k] = m[k];
    }
  }
  void set_state(Domain& d, int kind, const double* in) override {
    if (kind != XG_STATE_J2) return;
    const long long n = np();
    std::vector<double> e((size_t)ncomp() * n), x(n);
    for (long long p = 0; p < n; p++) {
      const double* m = in + p * 10 + 1;
      x[p] = in[p * 10];
      const double t[6] = {m[0], m[4], m[8], m[1], m[5], m[2]};
      if (!brick && (t[4] != 0.0 || t[5] != 0.0))
        throw Error(XG_ERR_ARG, "set_state: plane-strain points have no 12/20 plastic strain");
      for (int c = 0; c < ncomp(); c++) e[(size_t)c * n + p] = t[c];
    }
    epn.upload(e, d.stream);
    ept.upload(e, d.stream);
    xin.upload(x, d.stream);
    xit.upload(x, d.stream);
    XG_HIP(hipStreamSynchronize(d.stream));
  }
};

// ---------------------------------------------------------------------------
// Domain
// ---------------------------------------------------------------------------
Domain::~Domain() {
  cg_graph_clear();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  groups.clear();
  if (stream) (void)hipStreamDestroy(stream);
}

int Domain::lanes() {
  if (spmv_lanes > 0) return spmv_lanes;
  const double mean = num.neq ? (double)num.colind.size() / num.neq : 1.0;
  if (mean > 128) return 16;
  if (mean > 24) return 8;
  return 4;
}

// Tables of k_brick_tangent_csr for a model of stdBrick elements alone whose
// nodes have all 3 DOFs free or none (the node-blocked CSR view exists):
// per node block row I, the elements containing its node in ascending tag
// order (the group's order), the node's local index in each, and for each
// element node the position of its block in row I (255 for a fixed node).
// Off when another group is present, a row has more than kBrickCsrMaxNb
// neighbour blocks, or XG_BRICK_TANGENT_CSR=0 is set (the element-matrix
// path, for comparison).
void Domain::fused_tangent_tables() {
  fz_ok = false;
  fz_ptr.release();
  fz_ele.release();
  fz_a.release();
  fz_tgt.release();
  const char* env = std::getenv("XG_BRICK_TANGENT_CSR");
  if (env && std::strcmp(env, "0") == 0) return;
  if (groups.size() != 1 || nb_m != 3 || ndf != 3 || num.neq == 0) return;
  auto* jg = dynamic_cast<J2Group*>(groups[0].get());
  if (!jg || !jg->brick) return;
  const int nbr = num.neq / 3;
  const std::vector<int>& rp = num.rowptr;
  const std::vector<int>& ci = num.colind;
  std::vector<int> node_of(nbr, -1);
  for (int n = 0; n < nnode; n++)
    if (num.eq[(size_t)n * 3] >= 0) node_of[num.eq[(size_t)n * 3] / 3] = n;
  std::vector<int> cnt(nnode + 1, 0);
  const int ne = jg->nele();
  for (int e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++) cnt[jg->conn[8 * (size_t)e + a] + 1]++;
  for (int n = 0; n < nnode; n++) cnt[n + 1] += cnt[n];
  std::vector<int> ele_at(cnt[nnode]), loc_at(cnt[nnode]), fill(cnt.begin(), cnt.end() - 1);
  for (int e = 0; e < ne; e++)  // ascending e: each node's list in tag order
    for (int a = 0; a < 8; a++) {
      const int n = jg->conn[8 * (size_t)e + a];
      ele_at[fill[n]] = e;
      loc_at[fill[n]++] = a;
    }
  std::vector<int> ptr(nbr + 1, 0), ele;
  std::vector<unsigned char> loc, tgt;
  for (int I = 0; I < nbr; I++) {
    const int n = node_of[I];
    if (n < 0) return;
    const int b0 = rp[3 * I], len = rp[3 * I + 1] - b0, nnb = len / 3;
    if (nnb > kBrickCsrMaxNb) return;
    for (int k = cnt[n]; k < cnt[n + 1]; k++) {
      const int e = ele_at[k];
      ele.push_back(e);
      loc.push_back((unsigned char)loc_at[k]);
      for (int b = 0; b < 8; b++) {
        const int eqb = num.eq[(size_t)jg->conn[8 * (size_t)e + b] * 3];
        unsigned char t = 255;
        if (eqb >= 0) {
          const int* c0 = &ci[b0];
          const int* hit = std::lower_bound(c0, c0 + len, eqb);
          if (hit == c0 + len || *hit != eqb) return;  // not in the pattern: leave the element path on
          t = (unsigned char)((hit - c0) / 3);
        }
        tgt.push_back(t);
      }
    }
    ptr[I + 1] = (int)ele.size();
  }
  fz_ptr.upload(ptr, stream);
  fz_ele.upload(ele, stream);
  fz_a.upload(loc, stream);
  fz_tgt.upload(tgt, stream);
  fz_ok = true;
}

// Node-blocked view of the CSR pattern for k_spmv_nb_dot: with M = ndf, it
// exists when every node has all M DOFs free or none, its equations being
// M*J .. M*J + M-1, and every row of block J lists the same neighbour blocks
// as ascending runs of M columns.  Returns one block column per (block row,
// neighbour) and sets nb_m = M, or returns nothing and sets nb_m = 0.
std::vector<int> Domain::node_block_columns() {
  nb_m = 0;
  const int M = ndf;
  std::vector<int> out;
  if (M != 2 && M != 3 && M != 6) return out;
  for (int n = 0; n < nnode; n++) {
    const int* e = &num.eq[(size_t)n * M];
    if (e[0] < 0) {
      for (int j = 1; j < M; j++)
        if (e[j] >= 0) return out;
      continue;
    }
    if (e[0] % M) return out;
    for (int j = 1; j < M; j++)
      if (e[j] != e[0] + j) return out;
  }
  const int neq = num.neq;
  if (neq % M) return out;
  const std::vector<int>& rp = num.rowptr;
  const std::vector<int>& ci = num.colind;
  for (int J = 0; J < neq / M; J++) {
    const int r0 = M * J, b0 = rp[r0], len = rp[r0 + 1] - b0;
    if (len % M || (long long)b0 != (long long)M * M * (long long)out.size()) return {};
    for (int k = 0; k < len; k += M) {
      const int c0 = ci[b0 + k];
      if (c0 % M) return {};
      for (int c = 1; c < M; c++)
        if (ci[b0 + k + c] != c0 + c) return {};
      out.push_back(c0 / M);
    }
    for (int r = 1; r < M; r++) {
      const int b = rp[r0 + r];
      if (rp[r0 + r + 1] - b != len) return {};
      for (int k = 0; k < len; k++)
        if (ci[b + k] != ci[b0 + k]) return {};
    }
  }
  nb_m = M;
  return out;
}

// The block-symmetric copy of the matrix for the SpMV (XG_SPMV_SYM, and the
// PCG's products under XG_SPMV_AUTO; M = 2 or 3): the PCG assumes a
// symmetric matrix, so its products may read the upper blocks only.  Layout
// built on first use (sym_layout), stored blocks converted
// from the CSR values whenever an assembly has changed them.  sym_ok stays
// false where the block pattern is not symmetric, and the SpMV keeps the
// CSR values.
const SymSpmv* Domain::sym_prepare(bool solver) {
  const bool want = spmv_format == XG_SPMV_SYM || (solver && spmv_format == XG_SPMV_AUTO);
  if (!want || (nb_m != 2 && nb_m != 3) || num.neq == 0) return nullptr;
  if (!sym_built) {
    sym_built = true;
    sym_ok = false;
    const int M = nb_m, nbr = num.neq / M;
    const std::vector<int> nbc = node_block_columns();
    if (nb_m != M || nbc.empty()) return nullptr;
    // XG_SYM_ORDER=1 in the environment sweeps (and stores) the block rows in
    // the Morton order of their nodes (every node of a block row here has all
    // its DOFs free), the PCG's vectors in that order too: the partner reads
    // of the stored blocks then mostly hit L2 (64^3 cube: 1.15x instead of
    // 1.9x the algorithmic bytes fetched), but four consecutive block rows
    // share fewer cache lines than in the equation order, and the product is
    // bound by the L1's line lookups, not by HBM: 2-8% slower, so the
    // equation order is the default
    const char* ord_env = std::getenv("XG_SYM_ORDER");
    const bool use_ord = ord_env && ord_env[0] == '1' && nbr > 1;
    std::vector<int> ord;  // kept until the uploads have finished
    if (use_ord) {
      std::vector<double> xyz((size_t)nbr * ndm, 0.0);
      for (int nd = 0; nd < nnode; nd++) {
        const int e0 = num.eq[(size_t)nd * M];
        if (e0 < 0) continue;
        for (int d = 0; d < ndm; d++) xyz[(size_t)(e0 / M) * ndm + d] = crd_h[(size_t)nd * ndm + d];
      }
      ord = morton_order(nbr, ndm, xyz);
    }
    SymLayout L;
    if (!sym_layout(M, nbr, num.rowptr, nbc, L, ord)) return nullptr;
    if (use_ord) {
      // columns as sweep positions: x is read in the sweep order
      std::vector<int> pos(nbr);
      for (int q = 0; q < nbr; q++) pos[ord[q]] = q;
      for (int& c : L.col) c = pos[c];
    }
    hipStream_t s = stream;
    sym_ptr.upload(L.ptr, s);
    sym_col.upload(L.col, s);
    sym_blk.upload(L.blk, s);
    sym_src.upload(L.src, s);
    sym_len.upload(L.len, s);
    sym_sweep = use_ord;
    if (use_ord) {
      const int bwj = (system == XG_SYSTEM_PCG_BLOCKJACOBI) ? ndf : 1;
      sym_ord.upload(ord, s);
      sw_b.resize(num.neq);
      sw_x.resize(num.neq);
      binv_s.resize((size_t)num.neq * bwj);
    }
    sym_val.resize((size_t)L.nup * M * M);
    XG_HIP(hipStreamSynchronize(s));
    sym_nup = L.nup;
    sym_view.m = M;
    sym_view.nbr = nbr;
    sym_view.ptr = sym_ptr.p;
    sym_view.col = sym_col.p;
    sym_view.blk = sym_blk.p;
    sym_view.val = sym_val.p;
    sym_ok = true;
    sym_stale = true;
  }
  if (!sym_ok) return nullptr;
  if (spmv_lanes > 0) {
    sym_view.lanes = spmv_lanes;
  } else {
    // two passes over a mean block row (measured best on hexahedral meshes)
    const double half = 0.5 * (double)sym_col.n / std::max(sym_view.nbr, 1);
    sym_view.lanes = half > 32 ? 64 : half > 16 ? 32 : half > 8 ? 16 : half > 4 ? 8 : 4;
  }
  if (sym_stale) {
    launch_sym_convert(stream, sym_view, sym_nup, sym_src.p, sym_len.p, vals.p, sym_val.p);
    sym_stale = false;
  }
  return &sym_view;
}

void Domain::spmv(const double* x, double* y, bool solver) {
  const SymSpmv* sym = sym_prepare(solver);
  if (sym && sym_sweep) {
    // y = K x in the equation order through the product's sweep order
    launch_sweep_perm(stream, sym->nbr, sym->m, sym_ord.p, x, sw_b.p, false);
    spmv_sweep(sw_b.p, sw_x.p);
    launch_sweep_perm(stream, sym->nbr, sym->m, sym_ord.p, sw_x.p, y, true);
    return;
  }
  launch_spmv(stream, lanes(), num.neq, rowptr.p, colind.p, vals.p, x, y, nullptr, nullptr, spmv_nbm(), nbcol.p, sym);
}

void Domain::spmv_sweep(const double* x, double* y) {
  launch_spmv(stream, lanes(), num.neq, rowptr.p, colind.p, vals.p, x, y, nullptr, nullptr, spmv_nbm(), nbcol.p,
              sym_prepare(true));
}

void Domain::setup() {
  XG_HIP(hipSetDevice(device));
  cg_graph_clear();  // captured with the buffers of an earlier set-up
  if (nodedefs.empty()) throw Error(XG_ERR_STATE, "no nodes defined");
  // nodes in ascending tag order
  nnode = (int)nodedefs.size();
  node_tags.clear();
  node_index.clear();
  crd_h.assign((size_t)nnode * ndm, 0.0);
  std::vector<char> fixed((size_t)nnode * ndf, 0);
  std::vector<double> Pnode((size_t)nnode * ndf, 0.0);
  int i = 0;
  for (auto& kv : nodedefs) {
    node_tags.push_back(kv.first);
    node_index[kv.first] = i;
    for (int j = 0; j < ndm; j++) crd_h[(size_t)i * ndm + j] = kv.second.crd[j];
    for (int j = 0; j < ndf; j++) {
      fixed[(size_t)i * ndf + j] = kv.second.fix[j] ? 1 : 0;
      Pnode[(size_t)i * ndf + j] = kv.second.P[j];
    }
    i++;
  }
  // element groups in ascending tag order
  groups.clear();
  ele_order.clear();
  std::map<int, int> group_of_type;
  std::vector<std::vector<int>> elem_nodes;
  elem_nodes.reserve(eledefs.size());
  std::map<int, int> j2index;
  {
    int k = 0;
    for (auto& kv : j2defs) j2index[kv.first] = k++;
  }
  for (auto& kv : eledefs) {
    const EleDef& ed = kv.second;
    int gi;
    auto it = group_of_type.find(ed.type);
    if (it == group_of_type.end()) {
      std::unique_ptr<Group> g;
      if (ed.type == 1 || ed.type == 2) {
        auto jg = std::make_unique<J2Group>();
        jg->brick = ed.type == 1;
        jg->nen = jg->brick ? 8 : 4;
        jg->nd = jg->nen * ndf;
        g = std::move(jg);
      } else {
        g = make_frame_group(*this, ed.type == 4);
      }
      gi = (int)groups.size();
      groups.push_back(std::move(g));
      group_of_type[ed.type] = gi;
    } else
      gi = it->second;
    Group& g = *groups[gi];
    std::vector<int> idx;
    for (int nt : ed.nodes) {
      auto ni = node_index.find(nt);
      if (ni == node_index.end()) throw Error(XG_ERR_ARG, "element " + std::to_string(kv.first) + ": node not found");
      idx.push_back(ni->second);
    }
    ele_order.push_back({gi, g.nele()});
    g.tags.push_back(kv.first);
    g.conn.insert(g.conn.end(), idx.begin(), idx.end());
    if (ed.type == 1 || ed.type == 2) {
      J2Group& jg = static_cast<J2Group&>(g);
      jg.mat_h.push_back(j2index.at(ed.mat));
      if (jg.brick) jg.b_h.insert(jg.b_h.end(), {ed.b[0], ed.b[1], ed.b[2]});
      else {
        jg.b_h.insert(jg.b_h.end(), {ed.b[0], ed.b[1]});
        jg.thk_h.push_back(ed.thk);
      }
    } else {
      frame_group_add(*this, g, kv.first, ed);
    }
    elem_nodes.push_back(std::move(idx));
  }
  // numbering + CSR pattern
  number_and_pattern(nnode, ndf, fixed, elem_nodes, numberer == XG_NUMBERER_RCM, num);
  const int neq = num.neq;
  const long long nnz = (long long)num.colind.size();
  // element matrix / vector offsets
  ke_total = re_total = 0;
  for (auto& g : groups) {
    g->ke_off = ke_total;
    g->re_off = re_total;
    ke_total += (long long)g->nele() * g->nd * g->nd;
    re_total += (long long)g->nele() * g->nd;
  }
  if (ke_total > 2147483647LL) throw Error(XG_ERR_ARG, "element matrices exceed 2^31 entries");
  // gather maps, contributions in ascending element-tag order
  std::vector<int> cnt(nnz + 1, 0), rcnt(neq + 1, 0);
  std::vector<int> id;
  auto element_id = [&](size_t k, std::vector<int>& out) {
    out.clear();
    for (int n : elem_nodes[k])
      for (int j = 0; j < ndf; j++) out.push_back(num.eq[(size_t)n * ndf + j]);
  };
  auto slot = [&](int row, int col) -> long long {
    const int* b = num.colind.data() + num.rowptr[row];
    const int* e = num.colind.data() + num.rowptr[row + 1];
    const int* f = std::lower_bound(b, e, col);
    return f - num.colind.data();
  };
  std::vector<int> aidx, ridx, fillp, rfill;
  for (int pass = 0; pass < 2; pass++) {
    for (size_t k = 0; k < elem_nodes.size(); k++) {
      element_id(k, id);
      const Group& g = *groups[ele_order[k].first];
      const long long kbase = g.ke_off + (long long)ele_order[k].second * g.nd * g.nd;
      const long long rbase = g.re_off + (long long)ele_order[k].second * g.nd;
      const int nd = (int)id.size();
      for (int a = 0; a < nd; a++) {
        const int row = id[a];
        if (row < 0) continue;
        if (pass == 0) rcnt[row + 1]++;
        else ridx[rfill[row]++] = (int)(rbase + a);
        for (int b = 0; b < nd; b++) {
          const int col = id[b];
          if (col < 0) continue;
          const long long sl = slot(row, col);
          if (pass == 0) cnt[sl + 1]++;
          else aidx[fillp[sl]++] = (int)(kbase + (long long)a * nd + b);
        }
      }
    }
    if (pass == 0) {
      long long tot = 0;
      for (long long s2 = 0; s2 < nnz; s2++) {
        tot += cnt[s2 + 1];
        if (tot > 2147483647LL) throw Error(XG_ERR_ARG, "assembly map exceeds 2^31 entries");
        cnt[s2 + 1] = (int)tot;
      }
      for (int a = 0; a < neq; a++) rcnt[a + 1] += rcnt[a];
      aidx.resize(cnt[nnz]);
      ridx.resize(rcnt[neq]);
      fillp.assign(cnt.begin(), cnt.end() - 1);
      rfill.assign(rcnt.begin(), rcnt.end() - 1);
    }
  }
  fillp.clear();
  fillp.shrink_to_fit();
  // reference loads per equation (DOF-group order is irrelevant per entry)
  std::vector<double> Pe(neq, 0.0);
  for (int n = 0; n < nnode; n++)
    for (int j = 0; j < ndf; j++) {
      const int a = num.eq[(size_t)n * ndf + j];
      if (a >= 0) Pe[a] = Pnode[(size_t)n * ndf + j];
    }
  // nodal blocks for the block-Jacobi preconditioner
  std::vector<int> bf(neq, 0), blkf;
  std::vector<unsigned char> bs(neq, 1), blks;
  for (int n = 0; n < nnode; n++) {
    int first = -1, m = 0;
    for (int j = 0; j < ndf; j++) {
      const int a = num.eq[(size_t)n * ndf + j];
      if (a < 0) continue;
      if (first < 0) first = a;
      m++;
    }
    if (m == 0) continue;
    blkf.push_back(first);
    blks.push_back((unsigned char)m);
    for (int j = 0; j < m; j++) {
      bf[first + j] = first;
      bs[first + j] = (unsigned char)m;
    }
  }
  nblk = (int)blkf.size();
  std::vector<int> nbc = node_block_columns();

  // ---- device upload ----
  if (!stream) XG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (!ev0) {
    XG_HIP(hipEventCreate(&ev0));
    XG_HIP(hipEventCreate(&ev1));
  }
  hipStream_t s = stream;
  crd.upload(crd_h, s);
  U.resize((size_t)nnode * ndf);
  U.zero(s);
  Uc.resize(U.n);
  Uc.zero(s);
  V.resize(U.n);
  V.zero(s);
  Vc.resize(U.n);
  Vc.zero(s);
  A.resize(U.n);
  A.zero(s);
  Ac.resize(U.n);
  Ac.zero(s);
  eqmap.upload(num.eq, s);
  rowptr.upload(num.rowptr, s);
  colind.upload(num.colind, s);
  nbcol.upload(nbc, s);
  sym_built = sym_ok = false;
  sym_stale = true;
  sym_view = SymSpmv{};
  sym_sweep = false;
  for (auto* v : {&sym_ptr, &sym_col, &sym_blk, &sym_src, &sym_len, &sym_ord}) v->release();
  for (auto* v : {&sym_val, &sw_b, &sw_x, &binv_s}) v->release();
  vals.resize(nnz);
  vals.zero(s);
  Ke.resize(ke_total);
  Re.resize(re_total);
  asm_ptr.upload(cnt, s);
  asm_idx.upload(aidx, s);
  rhs_ptr.upload(rcnt, s);
  rhs_idx.upload(ridx, s);
  B.resize(neq);
  B.zero(s);
  X.resize(neq);
  Xw.resize(neq);
  xw_valid = false;
  dc_hat_valid = false;
  X.zero(s);
  Peq.upload(Pe, s);
  nfail.resize(1);
  nfail.zero(s);
  r0.resize(neq);
  r1.resize(neq);
  z.resize(neq);
  p.resize(neq);
  p2.resize(neq);
  q.resize(neq);
  tmpv.resize(neq);
  part.resize(cg_part_size(neq));
  pcbad.resize(1);
  binv.resize((size_t)neq * ndf);
  bfirst.upload(bf, s);
  bsize.upload(bs, s);
  blk_first.upload(blkf, s);
  blk_size.upload(blks, s);
  scal.resize(cg_scal_bytes());
  dscalar.resize(4);
  for (auto& g : groups) g->setup(*this);
  fused_tangent_tables();
  setup_transient();
  // the PCG's block-symmetric layout depends only on the pattern: built here
  // with the other index maps rather than in the first solve
  if (spmv_format != XG_SPMV_CSR) sym_prepare(true);
  band_order();
  pc_done = nullptr;
  wb_failed = false;
  XG_HIP(hipStreamSynchronize(s));
  lambda = 0;
  time = 0;
  dc_ready = false;
  ray_conn.clear();
  Kc.release();
  ready = true;
  rayleigh_kc_init();
}

// Per-equation nodal mass, influence vector of the UniformExcitation
// direction, equation -> node DOF map and the CSR positions of massive
// diagonals.
void Domain::setup_transient() {
  const int neq = num.neq;
  std::vector<double> m(neq, 0.0), r(neq, 0.0), mv;
  std::vector<int> e2d(neq, 0);
  std::vector<long long> pos;
  for (int n = 0; n < nnode; n++) {
    const NodeDef& nd = nodedefs.at(node_tags[n]);
    for (int j = 0; j < ndf; j++) {
      const int a = num.eq[(size_t)n * ndf + j];
      if (a < 0) continue;
      e2d[a] = n * ndf + j;
      m[a] = nd.mass[j];
      r[a] = (j == gm_dir) ? 1.0 : 0.0;
    }
  }
  for (int a = 0; a < neq; a++)
    if (m[a] != 0.0) {
      const int* b = num.colind.data() + num.rowptr[a];
      const int* e = num.colind.data() + num.rowptr[a + 1];
      pos.push_back(std::lower_bound(b, e, a) - num.colind.data());
      mv.push_back(m[a]);
    }
  nmass_eq = (int)pos.size();
  Meq.upload(m, stream);
  Req.upload(r, stream);
  eq2dof.upload(e2d, stream);
  mass_pos.upload(pos, stream);
  mass_val.upload(mv, stream);
}

// PathSeries::getFactor (PathSeries.cpp:334-356) scaled by the motion factor
double Domain::ground_accel(double t) const {
  if (gm_accel.empty() || gm_dt <= 0.0 || t < 0.0) return 0.0;
  const double incr = t / gm_dt;
  const int i1 = (int)std::floor(incr);
  const int i2 = i1 + 1;
  if (i2 >= (int)gm_accel.size()) return 0.0;
  const double v1 = gm_accel[i1], v2 = gm_accel[i2];
  return gm_factor * (v1 + (v2 - v1) * (incr - i1));
}

void Domain::update() {
  hipStream_t s = stream;
  XG_HIP(hipEventRecord(ev0, s));
  for (auto& g : groups) g->update(*this);
  XG_HIP(hipEventRecord(ev1, s));
  counters.updates++;
  for (auto& g : groups) counters.material_points += g->fibers();
}

// AnalysisModel::updateDomain returns Domain::update, the sum of the
// elements' update codes: here, whether an element group flagged a failure.
int Domain::update_checked() {
  update();
  bool bad = false;
  for (auto& g : groups) bad = g->update_failed(*this) || bad;
  return bad ? -1 : 0;
}

void Domain::form_unbalance() {
  for (auto& g : groups) g->resid(*this, Re.p + g->re_off);
  if (stiffness_damping()) rayleigh_forces();
  if (integ == 3)
    launch_assemble_rhs_dyn(stream, num.neq, rhs_ptr.p, rhs_idx.p, Re.p, lambda, Peq.p, Meq.p, Req.p, ag, alphaM,
                            eq2dof.p, V.p, A.p, B.p);
  else
    launch_assemble_rhs(stream, num.neq, rhs_ptr.p, rhs_idx.p, Re.p, lambda, Peq.p, B.p);
}

// Element::getRayleighDampingForces (Element.cpp:283-323), added to each
// element's resisting force by getResistingForceIncInertia: (betaK K +
// betaKc Kc) v with K the tangent of the current trial state.
void Domain::rayleigh_forces() {
  if (betaK != 0.0)
    for (auto& g : groups) g->tangent(*this, Ke.p + g->ke_off);
  if (ray_conn.size() != groups.size()) {
    ray_conn.clear();
    for (auto& g : groups) {
      ray_conn.emplace_back(new DVec<int>());
      ray_conn.back()->upload(g->conn, stream);
    }
  }
  for (size_t k = 0; k < groups.size(); k++) {
    Group& g = *groups[k];
    if (g.nd != g.nen * ndf) throw Error(XG_ERR_ARG, "rayleigh: element DOFs differ from the nodes' DOFs");
    launch_rayleigh_force(stream, g.nele(), g.nen, ndf, ray_conn[k]->p, Ke.p + g.ke_off,
                          betaKc != 0.0 ? Kc.p + g.ke_off : nullptr, betaK, betaKc, V.p, Re.p + g.re_off);
  }
}

// Element::setRayleighDampingFactors (Element.cpp:167-176): with betaKc the
// committed tangent starts as the tangent of the current state.
void Domain::rayleigh_kc_init() {
  if (betaKc == 0.0) {
    Kc.release();
    return;
  }
  if (Kc.n == (size_t)ke_total && Kc.p) return;
  Kc.resize(ke_total);
  for (auto& g : groups) g->tangent(*this, Kc.p + g->ke_off);
  XG_HIP(hipStreamSynchronize(stream));
}

void Domain::form_tangent() {
  // a stdBrick-only model without stiffness-proportional damping: element
  // tangents formed and summed into the CSR values in one pass
  const bool fused = groups.size() == 1 && !stiffness_damping() && groups[0]->tangent_csr(*this);
  if (!fused) {
    for (auto& g : groups) g->tangent(*this, Ke.p + g->ke_off);
    // Newmark::formEleTangent: K + c2 (betaK K + betaKc Kc) per element entry
    if (stiffness_damping()) launch_ke_damp(stream, ke_total, nm_c2, betaK, betaKc, Kc.p, Ke.p);
    launch_assemble_csr(stream, (long long)num.colind.size(), asm_ptr.p, asm_idx.p, Ke.p, vals.p);
  }
  sym_stale = true;
  pc_done = nullptr;
  wb_failed = false;
  // Newmark::formNodTangent: c2 C + c3 M of the nodal masses (C = alphaM M)
  if (integ == 3) launch_add_diag(stream, nmass_eq, mass_pos.p, mass_val.p, nm_c3 + nm_c2 * alphaM, vals.p);
}

Domain::CgBufs Domain::cg_bufs() {
  return CgBufs{r0.p, r1.p, z.p, p.p, p2.p, q.p, part.p, binv.p, pcbad.p, scal.p, band_A.p, band_L.p, band_Lr.p,
                sw_b.p, sw_x.p, binv_s.p, true};
}

// The preconditioner of the current matrix into w: the band-Cholesky factor
// for a system the one-workgroup band solve takes, else the (block-)Jacobi
// inverses.  Formed once per matrix and buffer set.
void Domain::precond_setup(const CgBufs& w) {
  const bool band = band_active();
  const bool wide = !band && !small_path() && w.wide && wide_active();
  const bool sweep = !band && !wide && !small_path() && sweep_active(w);
  const void* key = band   ? (const void*)w.bandL
                    : wide ? (const void*)wb_L.p
                    : sweep ? (const void*)w.binvs
                            : (const void*)w.binv;
  if (pc_done == key) return;
  // the domain's own factor storage (the allocation-free solve's narrow band
  // factor lives in its scratch instead)
  const bool own = (band && w.bandL == band_L.p) || wide;
  if (own && pc_cache && pc_cache_key == key) {
    XG_HIP(hipMemsetAsync(pc_diff.p, 0, sizeof(int), stream));
    launch_vals_differ(stream, (long long)num.colind.size(), vals.p, pc_vals.p, pc_diff.p);
    int diff = 1;
    XG_HIP(hipMemcpyAsync(&diff, pc_diff.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    XG_HIP(hipStreamSynchronize(stream));
    if (!diff) {  // the same matrix: its factor is in place (and had no bad pivot)
      XG_HIP(hipMemsetAsync(w.pcbad, 0, sizeof(int), stream));
      pc_reuses++;
      pc_done = key;
      return;
    }
  }
  if (own) {
    pc_cache = false;
    pc_fresh = pc_reuse_on && pc_vals.n >= num.colind.size();
    pc_fresh_key = key;
  }
  XG_HIP(hipMemsetAsync(w.pcbad, 0, sizeof(int), stream));
  if (wide) {
    wb_factor(w.pcbad);
  } else if (band) {
    band_factor(stream, num.neq, band_bw, rowptr.p, colind.p, vals.p, band_pinv.p, w.bandA, w.bandL, w.bandLr,
                w.pcbad);
  } else {
    const int bw = (system == XG_SYSTEM_PCG_BLOCKJACOBI) ? ndf : 1;
    launch_block_inverse(stream, nblk, blk_first.p, blk_size.p, rowptr.p, colind.p, vals.p, bw, w.binv, w.pcbad);
    // a node's M rows of bw entries each, to the sweep order
    if (sweep) launch_sweep_perm(stream, sym_view.nbr, sym_view.m * bw, sym_ord.p, w.binv, w.binvs, false);
  }
  pc_done = key;
}

// The wide-band factor of the current matrix; the launch sequence (one
// per tile column) replayed from a captured graph when cg_graph_on.
void Domain::wb_factor(int* bad) {
  const int n = num.neq;
  auto run = [&] {
    wband_factor(stream, n, wb_bw, rowptr.p, colind.p, vals.p, band_pinv.p, wb_L.p, wb_D.p, bad);
  };
  // plain launches when graphs are off or the flag is not the domain's own
  // (the allocation-free solve: the captured graph is left as it is)
  if (!cg_graph_on || bad != pcbad.p) {
    run();
    return;
  }
  if (!wb_fe) capture_graph(stream, run, &wb_fg, &wb_fe);
  XG_HIP(hipGraphLaunch(wb_fe, stream));
}

// z = M^-1 r for k = 1 or 2 right-hand sides, M the wide-band factor.
void Domain::wb_apply(int k, const double* const* r, double* const* z, bool graphs) {
  const int n = num.neq;
  const size_t np = (size_t)wband_padded(n);
  double* Yb = wb_Y.p;
  double* Yo = Yb + 2 * np;
  double* Xo = Yo + 2 * np;
  for (int s = 0; s < k; s++) wband_perm_in(stream, n, wb_bw, band_pinv.p, r[s], Yb + s * np);
  auto run = [&] { wband_sweeps(stream, n, wb_bw, k, wb_L.p, wb_D.p, Yb, Yo, Xo); };
  if (graphs) {
    if (!wb_se[k - 1]) capture_graph(stream, run, &wb_sg[k - 1], &wb_se[k - 1]);
    XG_HIP(hipGraphLaunch(wb_se[k - 1], stream));
  } else {
    run();
  }
  for (int s = 0; s < k; s++) wband_perm_out(stream, n, band_pinv.p, Xo + s * np, z[s]);
}

// PCG preconditioned by the exact wide-band factor, for R = 1 or 2
// right-hand sides side by side (the preconditioner applied to both in one
// set of sweeps): from x = 0, stopping at |r| <= rtol |b| like the other
// paths.  The factor makes the first iterate the direct solution up to
// rounding, so a solve takes one or two iterations.
int Domain::solve_wide(int R, const double* const* b, double* const* x, const CgBufs& w, bool graphs,
                       unsigned warm) {
  const int n = num.neq;
  const double tol2 = rtol * rtol;
  double* vr[2] = {w.r0, wb_v.p};
  double* vz[2] = {w.z, wb_v.p + n};
  double* vp[2] = {w.p, wb_v.p + 2 * (size_t)n};
  double* vq[2] = {w.q, wb_v.p + 3 * (size_t)n};
  double bb[2] = {0, 0}, rz[2] = {0, 0}, rr[2] = {0, 0};
  bool conv[2] = {true, true}, fail[2] = {false, false};
  int its[2] = {0, 0};
  for (int s = 0; s < R; s++) {
    launch_copy(stream, n, b[s], vr[s]);
    bb[s] = dot_dev(b[s], b[s], n);
    conv[s] = !(bb[s] > 0);
    if (!((warm >> s) & 1u) || conv[s]) {
      wband_zero(stream, n, x[s]);
    } else {
      // from the guess in x (bit s of warm): r = b - A x; a guess already
      // within the tolerance ends the solve before the first iteration
      spmv(x[s], vq[s], true);
      launch_axpby(stream, n, -1.0, vq[s], 1.0, vr[s]);
      rr[s] = dot_dev(vr[s], vr[s], n);
      conv[s] = rr[s] <= tol2 * bb[s];
    }
  }
  auto apply_open = [&] {
    const double* rs[2];
    double* zs[2];
    int k = 0;
    for (int s = 0; s < R; s++)
      if (!conv[s]) {
        rs[k] = vr[s];
        zs[k] = vz[s];
        k++;
      }
    if (k) wb_apply(k, rs, zs, graphs);
  };
  apply_open();
  for (int s = 0; s < R; s++)
    if (!conv[s]) {
      rz[s] = dot_dev(vr[s], vz[s], n);
      launch_copy(stream, n, vz[s], vp[s]);
    }
  for (int it = 0; it < maxit && !(conv[0] && conv[1]); it++) {
    for (int s = 0; s < R; s++) {
      if (conv[s]) continue;
      its[s]++;
      spmv(vp[s], vq[s], true);
      const double pq = dot_dev(vp[s], vq[s], n);
      if (pq == 0.0 || !std::isfinite(pq)) {
        conv[s] = fail[s] = true;
        continue;
      }
      const double alpha = rz[s] / pq;
      launch_axpby(stream, n, alpha, vp[s], 1.0, x[s]);
      launch_axpby(stream, n, -alpha, vq[s], 1.0, vr[s]);
      rr[s] = dot_dev(vr[s], vr[s], n);
      if (rr[s] <= tol2 * bb[s]) conv[s] = true;
    }
    if (conv[0] && conv[1]) break;
    apply_open();
    for (int s = 0; s < R; s++) {
      if (conv[s]) continue;
      const double rzn = dot_dev(vr[s], vz[s], n);
      const double beta = rzn / rz[s];
      rz[s] = rzn;
      launch_axpby(stream, n, 1.0, vz[s], beta, vp[s]);
    }
  }
  int status = 0;
  for (int s = 0; s < R; s++) {
    counters.solves++;
    counters.cg_iters += its[s];
    counters.last_cg_relres = bb[s] > 0 ? std::sqrt(rr[s] / bb[s]) : 0.0;
    if (fail[s] || (bb[s] > 0 && !(rr[s] <= tol2 * bb[s]))) status = XG_BAD_LINEAR_SOLVE;
  }
  return status;
}

// Half bandwidth of the pattern under an equation order (pinv: position of
// each equation).
static int band_width(const Numbering& num, const std::vector<int>& pinv) {
  int bw = 0;
  for (int r = 0; r < num.neq; r++)
    for (int q = num.rowptr[r]; q < num.rowptr[r + 1]; q++)
      bw = std::max(bw, std::abs(pinv[r] - pinv[num.colind[q]]));
  return bw;
}

void Domain::band_order() {
  band_bw = -1;
  wb_bw = -1;
  pc_cache = pc_fresh = false;
  pc_vals.release();
  pc_diff.release();
  band_pinv.release();
  band_A.release();
  band2.release();
  band_L.release();
  band_Lr.release();
  for (auto* v : {&wb_L, &wb_D, &wb_Y, &wb_v}) v->release();
  const int n = num.neq;
  if (n == 0 || n > kWbMaxEqn) return;
  std::vector<int> ident(n);
  for (int i = 0; i < n; i++) ident[i] = i;
  int best = band_width(num, ident);
  std::vector<int> pinv = ident;
  // nodes sorted along the longest extent of the model, then the other
  // axes, then DOF: the band of a slender mesh is its short side
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < nnode; i++)
    for (int k = 0; k < ndm; k++) {
      lo[k] = std::min(lo[k], crd_h[(size_t)i * ndm + k]);
      hi[k] = std::max(hi[k], crd_h[(size_t)i * ndm + k]);
    }
  int ax[3] = {0, 1, 2};
  std::sort(ax, ax + ndm, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
  std::vector<int> nodes(nnode);
  for (int i = 0; i < nnode; i++) nodes[i] = i;
  std::stable_sort(nodes.begin(), nodes.end(), [&](int a, int b) {
    for (int k = 0; k < ndm; k++) {
      const double xa = crd_h[(size_t)a * ndm + ax[k]], xb = crd_h[(size_t)b * ndm + ax[k]];
      if (xa != xb) return xa < xb;
    }
    return false;
  });
  std::vector<int> geo(n, -1);
  int k = 0;
  for (int nd : nodes)
    for (int j = 0; j < ndf; j++) {
      const int e = num.eq[(size_t)nd * ndf + j];
      if (e >= 0) geo[e] = k++;
    }
  if (k == n) {
    const int g = band_width(num, geo);
    if (g < best) {
      best = g;
      pinv = geo;
    }
  }
  if (best > 63 || n > kBandMaxEqn) {
    // the tiled factor of xg_wband.hip, for systems the one-workgroup
    // solve does not take
    if (n <= 1024 || best > kWbMaxBw) return;
    const size_t np = (size_t)wband_padded(n);
    const size_t need = (wband_factor_doubles(n, best) + wband_diag_doubles(n) + 6 * np + 4 * (size_t)n) * 8;
    if (need > kWbMaxBytes) return;
    wb_bw = best;
    band_pinv.upload(pinv, stream);
    wb_L.resize(wband_factor_doubles(n, best));
    wb_D.resize(wband_diag_doubles(n));
    wb_Y.resize(6 * np);
    wb_v.resize(4 * (size_t)n);
    pc_buffers();
    return;
  }
  band_bw = best;
  band_pinv.upload(pinv, stream);
  band_A.resize((size_t)n * (best + 1));
  band2.resize((size_t)5 * n);
  scal2.resize(cg_scal_bytes());
  band_L.resize((size_t)n * (best + 1));
  band_Lr.resize((size_t)n * (best + 1));
  pc_buffers();
}

// The copy of the factored matrix's values (and the comparison flag), for
// keeping the band factor of an unchanged matrix; XG_PC_REUSE=0 in the
// environment at setup refactors every assembled matrix instead.
void Domain::pc_buffers() {
  const char* env = std::getenv("XG_PC_REUSE");
  pc_reuse_on = !(env && env[0] == '0');
  if (!pc_reuse_on) return;
  pc_vals.resize(std::max<size_t>(num.colind.size(), 1));
  pc_diff.resize(1);
}

// Destroys the captured chunks once the stream has drained: a replay still
// queued or running must not outlive its executable graph.
void Domain::cg_graph_clear() {
  for (auto& g : cg_graphs) destroy_graph(stream, g.graph, g.exec);
  cg_graphs.clear();
  destroy_graph(stream, wb_fg, wb_fe);
  for (int k = 0; k < 2; k++) destroy_graph(stream, wb_sg[k], wb_se[k]);
}

// One chunk of PCG iterations.  A chunk of an even length starting on an
// even iteration (all of them but a solve's last, shortened one) is
// replayed from a graph of its launches: one host call instead of three to
// five kernel launches per iteration, which bounds the iteration rate of
// small systems.  Graphs hold kernels only, one per chunk length and set of
// launch arguments (at most kCgGraphs; further ones are launched one by
// one).
void Domain::cg_chunk_run(int iters, int it0, int L, int bw, double* x, const CgBufs& w, bool graphs) {
  const int n = num.neq, nbm = spmv_nbm();
  const SymSpmv* sym = sym_prepare(true);
  // block inverses in the order the vectors are in
  const double* binvp = sweep_active(w) ? w.binvs : w.binv;
  auto eager = [&](hipStream_t s, int k, int i0) {
    cg_iterations(s, k, i0, L, n, rowptr.p, colind.p, vals.p, x, w.r0, w.r1, w.z, w.p, w.p2, w.q, bw, binvp,
                  bfirst.p, bsize.p, w.part, w.scal, nbm, nbcol.p, cg_fuse_on, sym);
  };
  if (!graphs || (iters & 1) || (it0 & 1)) {
    eager(stream, iters, it0);
    return;
  }
  const std::vector<uintptr_t> key = {
      (uintptr_t)iters,   (uintptr_t)L,       (uintptr_t)n,       (uintptr_t)bw,      (uintptr_t)nbm, (uintptr_t)cg_fuse_on,
      (uintptr_t)x,       (uintptr_t)rowptr.p, (uintptr_t)colind.p, (uintptr_t)vals.p, (uintptr_t)w.r0,
      (uintptr_t)w.r1,    (uintptr_t)w.z,     (uintptr_t)w.p,     (uintptr_t)w.p2,    (uintptr_t)w.q,     (uintptr_t)binvp,
      (uintptr_t)bfirst.p, (uintptr_t)bsize.p, (uintptr_t)w.part, (uintptr_t)w.scal,  (uintptr_t)nbcol.p,
      (uintptr_t)(sym ? sym->val : nullptr), (uintptr_t)(sym ? sym->lanes : 0)};
  for (auto& g : cg_graphs)
    if (g.key == key) {
      XG_HIP(hipGraphLaunch(g.exec, stream));
      return;
    }
  if ((int)cg_graphs.size() >= kCgGraphs) {
    eager(stream, iters, it0);
    return;
  }
  // the launch checks of cg_iterations run once outside the capture, so a
  // bad argument throws before the stream enters capture mode
  cg_iterations_check(bw, sym);
  CgGraph g;
  g.key = key;
  capture_graph(stream, [&] { eager(stream, iters, 0); }, &g.graph, &g.exec);
  cg_graphs.push_back(g);
  XG_HIP(hipGraphLaunch(g.exec, stream));
}

// Two solves with the current matrix, B1 -> X1 and B2 -> X2
// (DisplacementControl::update's unbalance and reference-load solves).  With
// the band preconditioner the two run side by side in one launch (one
// workgroup each, the same kernel and factor as solve(), so the same
// results); otherwise one after the other.
int Domain::solve_pair(const double* b1, double* x1, const double* b2, double* x2) {
  const int n = num.neq;
  if (n > 0 && !small_path() && wide_active()) {
    const CgBufs w = cg_bufs();
    if (!precond_ready(w)) return XG_BAD_LINEAR_SOLVE;
    if (wide_active()) {
      const double* bs[2] = {b1, b2};
      double* xs[2] = {x1, x2};
      // the second right-hand side is DisplacementControl's reference load:
      // its last solution in x2 is the start
      return solve_wide(2, bs, xs, w, cg_graph_on, (cg_warm_on && dc_hat_valid) ? 2u : 0u);
    }
    // the factor gave way to block Jacobi: one solve after the other
    const int r = solve(b1, x1);
    return r < 0 ? r : solve(b2, x2);
  }
  if (n == 0 || !band_active() || band2.n < (size_t)5 * n) {
    const int r = solve(b1, x1);
    return r < 0 ? r : solve(b2, x2);
  }
  const CgBufs w = cg_bufs();
  if (!precond_ready(w)) return XG_BAD_LINEAR_SOLVE;
  if (!band_active()) {  // the factor gave way to the PCG without it
    const int r = solve(b1, x1);
    return r < 0 ? r : solve(b2, x2);
  }
  const BandRhs R1{b1, x1, w.r0, w.z, w.p, w.q, w.r1, (CGScal*)w.scal};
  double* v = band2.p;
  const BandRhs R2{b2, x2, v, v + n, v + 2 * (size_t)n, v + 3 * (size_t)n, v + 4 * (size_t)n, (CGScal*)scal2.p};
  cg_solve_band(stream, n, rowptr.p, colind.p, vals.p, band_bw, band_pinv.p, w.bandL, w.bandLr, rtol * rtol, maxit,
                R1, &R2);
  int bad = 0;
  XG_HIP(hipMemcpyAsync(&bad, w.pcbad, sizeof(int), hipMemcpyDeviceToHost, stream));
  int status = 0;
  for (int k = 0; k < 2; k++) {
    int conv = 0, it = 0, brk = 0;
    double rr = 0, bb = 0;
    cg_read_scal(stream, k == 0 ? w.scal : (void*)scal2.p, &conv, &it, &rr, &bb, &brk);
    counters.solves++;
    counters.cg_iters += it;
    counters.last_cg_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (bb > 0 && !(rr <= rtol * rtol * bb)) status = XG_BAD_LINEAR_SOLVE;
  }
  if (bad) return XG_BAD_LINEAR_SOLVE;
  return status;
}

// The preconditioner of the current matrix into w, checked: a band factor
// that met a pivot that is not positive gives way to the PCG without it.
bool Domain::precond_ready(const CgBufs& w) {
  for (;;) {
    const bool wide = band_active() || (w.wide && wide_active() && !small_path());
    precond_setup(w);
    int bad = 0;
    XG_HIP(hipMemcpyAsync(&bad, w.pcbad, sizeof(int), hipMemcpyDeviceToHost, stream));
    XG_HIP(hipStreamSynchronize(stream));
    if (!bad) {
      if (pc_fresh) {  // a factor just formed in the domain's own storage: remember its matrix
        launch_copy_d(stream, (long long)num.colind.size(), vals.p, pc_vals.p);
        pc_cache = true;
        pc_cache_key = pc_fresh_key;
      }
      pc_fresh = false;
      return true;
    }
    pc_fresh = false;
    if (!wide) return false;
    wb_failed = true;
    wb_fallbacks++;
    pc_done = nullptr;
  }
}

int Domain::solve_with(const double* b, double* x, const CgBufs& w, bool graphs, bool warm) {
  const int n = num.neq;
  if (n == 0) return 0;
  const int bw = (system == XG_SYSTEM_PCG_BLOCKJACOBI) ? ndf : 1;
  if (!precond_ready(w)) return XG_BAD_LINEAR_SOLVE;
  int conv = 0, it = 0, brk = 0;
  double rr = 0, bb = 0;
  if (band_active()) {
    // one workgroup, one launch, preconditioned by the band's exact factor
    const BandRhs R{b, x, w.r0, w.z, w.p, w.q, w.r1, (CGScal*)w.scal, warm ? 1 : 0};
    cg_solve_band(stream, n, rowptr.p, colind.p, vals.p, band_bw, band_pinv.p, w.bandL, w.bandLr, rtol * rtol, maxit,
                  R);
    cg_read_scal(stream, w.scal, &conv, &it, &rr, &bb, &brk);
  } else if (!small_path() && w.wide && wide_active()) {
    const double* bs[1] = {b};
    double* xs[1] = {x};
    return solve_wide(1, bs, xs, w, graphs, warm ? 1u : 0u);
  } else if (small_path()) {
    // the whole solve in one workgroup, one launch
    cg_solve_small(stream, n, rowptr.p, colind.p, vals.p, b, x, w.r0, w.z, w.p, w.q, bw, w.binv, bfirst.p, bsize.p,
                   rtol * rtol, maxit, w.scal);
    cg_read_scal(stream, w.scal, &conv, &it, &rr, &bb, &brk);
  } else {
    // in the sweep order of the symmetric product: b (and a warm start's
    // guess) gathered into it here, the solution scattered back at the end
    const bool sweep = sweep_active(w);
    const double* bs = b;
    double* xs = x;
    if (sweep) {
      launch_sweep_perm(stream, sym_view.nbr, sym_view.m, sym_ord.p, b, w.bs, false);
      if (warm) launch_sweep_perm(stream, sym_view.nbr, sym_view.m, sym_ord.p, x, w.xs, false);
      bs = w.bs;
      xs = w.xs;
    }
    // a warm start takes the guess in x: q = A x first
    if (warm) spmv_sweep(xs, w.q);
    cg_start(stream, n, bs, xs, w.r0, w.z, w.p, bw, sweep ? w.binvs : w.binv, bfirst.p, bsize.p, w.part,
             rtol * rtol, maxit, w.scal, warm ? w.q : nullptr);
    int issued = 0, next = cg_chunk;
    const int L = lanes();
    while (true) {
      cg_read_scal(stream, w.scal, &conv, &it, &rr, &bb, &brk);
      if (conv) break;
      // maxit + 1 iterations at most: the fused iteration tests the previous
      // update's convergence at its start, so the last test comes one
      // iteration later (the extra iteration then returns at once, as it
      // does on the other path, where the test ends the iteration)
      if (issued >= maxit + 1) break;
      // chunks double (up to 8x cg_chunk) while the solve goes on: fewer
      // host round trips on long solves; iterations issued past
      // convergence return at once
      const int chunk = std::min(next, maxit + 1 - issued);
      cg_chunk_run(chunk, issued, L, bw, xs, w, graphs);
      issued += chunk;
      next = std::min(2 * next, 8 * cg_chunk);
    }
    if (sweep) launch_sweep_perm(stream, sym_view.nbr, sym_view.m, sym_ord.p, w.xs, x, true);
  }
  counters.solves++;
  counters.cg_iters += it;
  counters.last_cg_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  if (brk && bb > 0 && rr > rtol * rtol * bb) return XG_BAD_LINEAR_SOLVE;
  if (bb > 0 && !(rr <= rtol * rtol * bb)) return XG_BAD_LINEAR_SOLVE;
  return 0;
}

void Domain::incr_disp(const double* dUdev, double scale) {
  launch_incr_disp(stream, (long long)nnode * ndf, eqmap.p, dUdev, scale, U.p);
}

void Domain::commit() {
  // Element::commitState (Element.cpp:74-80) first: Kc = getTangentStiff()
  if (betaKc != 0.0 && Kc.p)
    for (auto& g : groups) g->tangent(*this, Kc.p + g->ke_off);
  for (auto& g : groups) g->commit(*this);
  Uc.copy_from(U, stream);
  Vc.copy_from(V, stream);
  Ac.copy_from(A, stream);
  XG_HIP(hipStreamSynchronize(stream));
}

// Domain::revertToLastCommit (Domain.cpp:1925-1939): nodes and elements
// revert, then one more update at the committed displacements with dT = 0.
void Domain::revert() {
  U.copy_from(Uc, stream);
  V.copy_from(Vc, stream);
  A.copy_from(Ac, stream);
  for (auto& g : groups) g->revert(*this);
  const double dt_keep = dt;
  dt = 0.0;
  if (ready) update();
  dt = dt_keep;
  XG_HIP(hipStreamSynchronize(stream));
}

// Domain::revertToStart (Domain.cpp:1941-1958): nodes and elements to their
// initial state, then one update at zero displacement with dT = 0.
void Domain::revert_to_start() {
  U.zero(stream);
  Uc.zero(stream);
  V.zero(stream);
  Vc.zero(stream);
  A.zero(stream);
  Ac.zero(stream);
  for (auto& g : groups) g->revert_to_start(*this);
  lambda = 0;
  time = 0;
  const double dt_keep = dt;
  dt = 0.0;
  if (ready) update();
  dt = dt_keep;
  XG_HIP(hipStreamSynchronize(stream));
}

double Domain::dot_dev(const double* a, const double* b, int n) {
  launch_dot(stream, n, a, b, part.p, dscalar.p);
  double h = 0;
  dscalar.download(&h, 1, stream);
  return h;
}
double Domain::norm2_dev(const double* v, int n) { return std::sqrt(dot_dev(v, v, n)); }
double Domain::read_eq(const double* v, int a) {
  double h = 0;
  XG_HIP(hipMemcpyAsync(&h, v + a, sizeof(double), hipMemcpyDeviceToHost, stream));
  XG_HIP(hipStreamSynchronize(stream));
  return h;
}

// ---- convergence tests ----------------------------------------------------
int Domain::run_test(int& iter, std::vector<double>& norms, int& numIncr) {
  double norm;
  const int n = num.neq;
  if (test_kind == XG_TEST_NORM_UNBALANCE) norm = norm2_dev(B.p, n);  // CTestNormUnbalance.cpp:81
  else if (test_kind == XG_TEST_NORM_DISP_INCR) norm = norm2_dev(X.p, n);  // CTestNormDispIncr.cpp:78
  else norm = 0.5 * std::fabs(dot_dev(X.p, B.p, n));  // CTestEnergyIncr.cpp:85
  counters.last_norm = norm;
  if (iter > 1 && !norms.empty() && norms.back() < norm) numIncr++;
  norms.push_back(norm);
  if (norm <= test_tol) return iter;
  // numIncr (norm increases) only fails CTestNormUnbalance (CTestNormUnbalance.cpp:172)
  if (iter >= test_maxit || (test_kind == XG_TEST_NORM_UNBALANCE && numIncr >= test_maxit)) {
    iter++;
    return -2;
  }
  iter++;
  return -1;
}

// ---- integrators --------------------------------------------------------------
int Domain::new_step() {
  const int n = num.neq;
  if (integ == 3) {  // Newmark::newStep, Newmark.cpp:105-170 (displacement formulation)
    const double g = nm_gamma, b = nm_beta, h = dt;
    nm_c2 = g / (b * h);
    nm_c3 = 1.0 / (b * h * h);
    const double a1 = (1.0 - g / b);
    const double a2 = h * (1.0 - 0.5 * g / b);
    const double a3 = -1.0 / (b * h);
    const double a4 = 1.0 - 0.5 / b;
    launch_newmark_predict(stream, (long long)nnode * ndf, eqmap.p, Uc.p, Vc.p, Ac.p, a1, a2, a3, a4, U.p, V.p, A.p);
    time += h;
    ag = ground_accel(time);  // EarthquakePattern::applyLoad(time)
    if (update_checked() < 0) return XG_BAD_STEP;
    return 0;
  }
  if (integ == 1) {  // LoadControl::newStep, LoadControl.cpp:70-95
    lc_dlambda *= lc_spec / lc_last;
    if (lc_dlambda < lc_min) lc_dlambda = lc_min;
    else if (lc_dlambda > lc_max) lc_dlambda = lc_max;
    lambda += lc_dlambda;
    lc_last = 0;
    return 0;
  }
  // DisplacementControl::newStep, DisplacementControl.cpp:121-222
  dc_incr *= dc_spec / dc_last;
  if (dc_incr < dc_min) dc_incr = dc_min;
  else if (dc_incr > dc_max) dc_incr = dc_max;
  form_tangent();
  if (solve_from(phat.p, dUhat.p, cg_warm_on && dc_hat_valid) < 0) return XG_BAD_STEP;
  dc_hat_valid = true;
  const double dUahat = read_eq(dUhat.p, dc_eq);
  if (dUahat == 0.0) return XG_BAD_STEP;
  const double dl = dc_incr / dUahat;
  lambda += dl;
  XG_HIP(hipMemsetAsync(dU.p, 0, n * sizeof(double), stream));
  launch_axpby(stream, n, dl, dUhat.p, 0.0, dU.p);
  incr_disp(dU.p);
  if (update_checked() < 0) return XG_BAD_STEP;
  dc_last = 0;
  return 0;
}

int Domain::integ_update(const double* x) {
  const int n = num.neq;
  if (integ == 3) {  // Newmark::update, Newmark.cpp:411-458
    launch_newmark_correct(stream, (long long)nnode * ndf, eqmap.p, x, nm_c2, nm_c3, U.p, V.p, A.p);
    if (update_checked() < 0) return XG_BAD_STEP;
    return 0;
  }
  if (integ == 1) {  // LoadControl::update, LoadControl.cpp:102-120
    incr_disp(x);
    if (update_checked() < 0) return XG_BAD_STEP;
    lc_last++;
    return 0;
  }
  // DisplacementControl::update, DisplacementControl.cpp:224-283
  XG_HIP(hipMemcpyAsync(dUbar.p, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  const double dUabar = read_eq(dUbar.p, dc_eq);
  const bool hat_ready = dc_hat_ready;
  dc_hat_ready = false;
  if (!hat_ready && solve_from(phat.p, dUhat.p, cg_warm_on && dc_hat_valid) < 0) return XG_BAD_STEP;
  dc_hat_valid = true;
  const double dUahat = read_eq(dUhat.p, dc_eq);
  if (dUahat == 0.0) return XG_BAD_STEP;
  const double dL = -dUabar / dUahat;
  XG_HIP(hipMemcpyAsync(dU.p, dUbar.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  launch_axpby(stream, n, dL, dUhat.p, 1.0, dU.p);
  lambda += dL;
  incr_disp(dU.p);
  if (update_checked() < 0) return XG_BAD_STEP;
  XG_HIP(hipMemcpyAsync(X.p, dU.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  dc_last++;
  return 0;
}

// NewtonRaphson::solveCurrentStep, NewtonRaphson.cpp:125-221
int Domain::newton() {
  form_unbalance();
  int iter = 1, numIncr = 0, result = -1;
  std::vector<double> norms;
  bool first = true;
  do {
    form_tangent();
    if (integ == 2 && (band_active() || (wide_active() && !small_path()))) {
      // DisplacementControl::update solves for the reference load with the
      // same matrix: both solves at once
      if (solve_pair(B.p, X.p, phat.p, dUhat.p) < 0) return XG_BAD_LINEAR_SOLVE;
      dc_hat_ready = true;
    } else {
      // under LoadControl the first iteration of a step solves for the same
      // load increment as the last step's first iteration did, on a tangent
      // that changed little: start it from that solution.  (The first
      // residual of a DisplacementControl or Newmark step is the unbalance
      // left by the predictor, which the last step's solve does not guess.)
      const bool warm_step = cg_warm_on && first && integ == 1;
      const bool warm = warm_step && xw_valid;
      if (warm) launch_copy_d(stream, num.neq, Xw.p, X.p);
      if (solve_from(B.p, X.p, warm) < 0) return XG_BAD_LINEAR_SOLVE;
      if (warm_step) {
        launch_copy_d(stream, num.neq, X.p, Xw.p);
        xw_valid = true;
      }
    }
    first = false;
    if (integ_update(X.p) < 0) return XG_BAD_STEP;
    form_unbalance();
    result = run_test(iter, norms, numIncr);
    counters.newton_iters++;
  } while (result == -1);
  if (result == -2) return XG_TEST_FAILED;
  return 0;
}

int Domain::analyze(int nsteps, double dtv) {
  if (!ready) setup();
  if (integ == 3) {
    if (!(dtv > 0.0)) throw Error(XG_ERR_ARG, "analyze: transient analysis needs dt > 0");
    for (auto& kv : eledefs)
      if (kv.second.rho != 0.0) throw Error(XG_ERR_ARG, "transient analysis: element mass density is not supported");
    dt = dtv;
  }
  if (integ == 2 && !dc_ready) {
    // DisplacementControl::domainChanged, :286-415: phat = B(lambda+1) at the
    // current (zero-unbalance) state
    const int n = num.neq;
    phat.resize(n);
    dUhat.resize(n);
    dUhat.zero(stream);
    dc_hat_valid = false;
    dUbar.resize(n);
    dU.resize(n);
    lambda += 1.0;
    form_unbalance();
    XG_HIP(hipMemcpyAsync(phat.p, B.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
    lambda -= 1.0;
    auto it = node_index.find(dc_node);
    if (it == node_index.end()) throw Error(XG_ERR_ARG, "DisplacementControl: node not found");
    dc_eq = num.eq[(size_t)it->second * ndf + dc_dof];
    if (dc_eq < 0) throw Error(XG_ERR_ARG, "DisplacementControl: dof is fixed");
    if (norm2_dev(phat.p, n) == 0.0) throw Error(XG_ERR_ARG, "DisplacementControl: zero reference load");
    dc_ready = true;
  }
  (void)dtv;
  for (int st = 0; st < nsteps; st++) {
    const double lam0 = lambda, t0 = time;
    int r = new_step();
    if (r == 0) r = newton();
    if (r < 0) {
      revert();
      lambda = lam0;
      time = t0;
      return r;
    }
    commit();
  }
  return 0;
}

// ---- committed state in the reference's sendSelf layouts --------------------
long long Domain::state_size(int kind) {
  if (!ready) setup();
  if (kind == XG_STATE_NODES) return (long long)nnode * ndf * 3;
  if (kind == XG_STATE_DOMAIN) return 2;
  if (kind != XG_STATE_J2 && kind != XG_STATE_STEEL02 && kind != XG_STATE_CONCRETE02)
    throw Error(XG_ERR_ARG, "state: unknown kind");
  long long n = 0;
  for (auto& g : groups) n += g->state_size(kind);
  return n;
}

// committed displacement, velocity, acceleration per node (Node::sendSelf,
// Node.cpp:767-791), nodes in ascending tag order
void Domain::get_state(int kind, double* out) {
  const long long size = state_size(kind);
  if (kind == XG_STATE_NODES) {
    const std::vector<double> u = Uc.download(stream), v = Vc.download(stream), a = Ac.download(stream);
    for (int n = 0; n < nnode; n++)
      for (int j = 0; j < ndf; j++) {
        out[(size_t)n * 3 * ndf + j] = u[(size_t)n * ndf + j];
        out[(size_t)n * 3 * ndf + ndf + j] = v[(size_t)n * ndf + j];
        out[(size_t)n * 3 * ndf + 2 * ndf + j] = a[(size_t)n * ndf + j];
      }
    return;
  }
  if (kind == XG_STATE_DOMAIN) {
    out[0] = time;
    out[1] = lambda;
    return;
  }
  long long off = 0;
  for (auto& g : groups) {
    g->get_state(*this, kind, out + off);
    off += g->state_size(kind);
  }
  (void)size;
}

void Domain::set_state(int kind, const double* in) {
  (void)state_size(kind);
  if (kind == XG_STATE_NODES) {
    std::vector<double> u((size_t)nnode * ndf), v(u.size()), a(u.size());
    for (int n = 0; n < nnode; n++)
      for (int j = 0; j < ndf; j++) {
        u[(size_t)n * ndf + j] = in[(size_t)n * 3 * ndf + j];
        v[(size_t)n * ndf + j] = in[(size_t)n * 3 * ndf + ndf + j];
        a[(size_t)n * ndf + j] = in[(size_t)n * 3 * ndf + 2 * ndf + j];
      }
    for (DVec<double>* t : {&U, &Uc}) t->upload(u, stream);
    for (DVec<double>* t : {&V, &Vc}) t->upload(v, stream);
    for (DVec<double>* t : {&A, &Ac}) t->upload(a, stream);
    XG_HIP(hipStreamSynchronize(stream));
    return;
  }
  if (kind == XG_STATE_DOMAIN) {
    time = in[0];
    lambda = in[1];
    return;
  }
  long long off = 0;
  for (auto& g : groups) {
    g->set_state(*this, kind, in + off);
    off += g->state_size(kind);
  }
}

}  // namespace xg
