// Oracle: CPU restatement of the reference Domain / analysis path:
// PlainHandler + RCM/Plain numbering + DOF graph + CSR setSize, element-order
// assembly, a profile LDL^T direct solve (the role of the reference default
// ProfileSPDLinSOE), NewtonRaphson with LoadControl / DisplacementControl /
// Newmark and the NormUnbalance / NormDispIncr / EnergyIncr tests.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU library:
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
// the product never does.  It exports the same C-ABI as include/xara_gpu.h
// with the prefix `orc_` instead of `xg_`.
#include "orc_ddm.hpp"
#include "orc_elements.hpp"
#include "orc_force.hpp"
#include "orc_frame.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/xara_gpu.h"

namespace orc {

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Element interface (the subset of SRC/element/Element.h:44-124 on the path)
// ---------------------------------------------------------------------------
struct Dom;
struct OElement {
  int tag = 0;
  std::vector<int> nodes;  // node indices (tag order positions)
  // committed tangent for betaKcomm damping (Element::Kc, Element.cpp:167-176):
  // formed when the factors are set, refreshed by every commitState (:74-80)
  std::vector<double> Kc;
  virtual ~OElement() {}
  virtual int ndof() const = 0;
  virtual int update(Dom& d) = 0;
  virtual void tangent(Dom& d, double* K) = 0;  // row-major ndof x ndof
  virtual void resist(Dom& d, double* R) = 0;
  virtual void mass(Dom& d, double* M) { std::fill(M, M + ndof() * ndof(), 0.0); }
  virtual void commit() = 0;
  virtual void revert() = 0;
  virtual void revert_to_start() = 0;
  virtual int npoints() const = 0;
  virtual int point_data(int what, double* out) const = 0;  // returns width
};

struct NodeRec {
  int tag;
  double crd[3] = {0, 0, 0};
  double U[6] = {0}, Uc[6] = {0};         // trial / committed displacement
  double V[6] = {0}, Vc[6] = {0}, A[6] = {0}, Ac[6] = {0};
  int fix[6] = {0};
  double P[6] = {0};
  double mass[6] = {0};
  int eq[6] = {-2, -2, -2, -2, -2, -2};
};

struct Dom {
  int ndm, ndf;
  std::map<int, NodeRec> nodemap;       // before setup
  std::vector<NodeRec> nodes;            // ascending tag after setup
  std::map<int, int> node_index;
  std::map<int, J2Params> j2;
  std::map<int, Steel02Params> steel;
  std::map<int, Concrete02Params> conc;
  std::map<int, double> elastic;
  std::map<int, FiberSectionDef> sections;
  std::map<int, std::vector<double>> transfs;
  std::map<int, std::unique_ptr<OElement>> elemap;
  std::vector<OElement*> eles;          // ascending tag after setup
  bool ready = false;
  // analysis
  int system = XG_SYSTEM_DENSE, numberer = XG_NUMBERER_RCM;
  double sys_rtol = 1e-12;
  int sys_maxit = 10000;
  int test_kind = XG_TEST_NORM_UNBALANCE;
  double test_tol = 1e-6;
  int test_maxit = 25;
  int integ = 1;  // 1 load control, 2 disp control, 3 newmark
  double lc_dlambda = 1.0, lc_dlmin = 1.0, lc_dlmax = 1.0, lc_spec = 1.0, lc_last = 1.0;
  int dc_node = 0, dc_dof = 0;
  double dc_incr = 0, dc_min = 0, dc_max = 0, dc_spec = 1.0, dc_last = 1.0;
  int dc_dofid = -1;
  std::vector<double> phat, dUhat, dUbar, dU, dUstep;
  double dc_dlambda_step = 0;
  double nm_gamma = 0.5, nm_beta = 0.25;
  double alphaM = 0, betaK = 0, betaKinit = 0, betaKcomm = 0;
  int gm_dir = -1;
  double gm_dt = 0, gm_factor = 1;
  std::vector<double> gm_accel;
  double lambda = 0, time = 0, dt = 0;
  // system
  int neq = 0;
  std::vector<int> rowptr, colind;
  std::vector<double> Aval, B, X;
  xg_counters counters{};
};

// ---------------------------------------------------------------------------
// Elements
// ---------------------------------------------------------------------------
struct OBrick : OElement {
  J2Params mp;
  double b[3];
  J2Point pts[8];
  int ndof() const override { return 24; }
  void gather(Dom& d, double xl[3][8], double ul[8][3]) {
    for (int i = 0; i < 8; i++) {
      const NodeRec& n = d.nodes[nodes[i]];
      for (int j = 0; j < 3; j++) {
        xl[j][i] = n.crd[j];
        ul[i][j] = n.U[j];
      }
    }
  }
  int update(Dom& d) override {
    double xl[3][8], ul[8][3];
    gather(d, xl, ul);
    brick_update(mp, d.dt, pts, xl, ul);
    return 0;  // Brick::update always returns 0 (Brick.cpp:839)
  }
  void tangent(Dom& d, double* K) override {
    double xl[3][8], ul[8][3], R[24];
    gather(d, xl, ul);
    brick_form_resid_and_tangent(pts, xl, b, 1, K, R);
  }
  void resist(Dom& d, double* R) override {
    double xl[3][8], ul[8][3];
    std::vector<double> K(576);
    gather(d, xl, ul);
    brick_form_resid_and_tangent(pts, xl, b, 0, K.data(), R);
  }
  void commit() override {
    for (auto& p : pts) j2_commit(p);
  }
  void revert() override {}  // J2 revertToLastCommit is a no-op
  void revert_to_start() override {
    for (auto& p : pts) p.zero();
  }
  int npoints() const override { return 8; }
  int point_data(int what, double* out) const override {
    for (int g = 0; g < 8; g++) {
      if (what == XG_OUT_STRESS) j2_3d_get_stress(pts[g], out + 6 * g);
      else if (what == XG_OUT_TANGENT) j2_3d_get_tangent(pts[g], out + 36 * g);
      else if (what == XG_OUT_STRAIN) {
        const J2Point& s = pts[g];
        double* o = out + 6 * g;
        o[0] = s.strain[0][0]; o[1] = s.strain[1][1]; o[2] = s.strain[2][2];
        o[3] = 2.0 * s.strain[0][1]; o[4] = 2.0 * s.strain[1][2]; o[5] = 2.0 * s.strain[2][0];
      } else if (what == XG_OUT_PLASTIC) {
        const J2Point& s = pts[g];
        double* o = out + 7 * g;
        o[0] = s.epsp_np1[0][0]; o[1] = s.epsp_np1[1][1]; o[2] = s.epsp_np1[2][2];
        o[3] = s.epsp_np1[0][1]; o[4] = s.epsp_np1[1][2]; o[5] = s.epsp_np1[2][0];
        o[6] = s.xi_np1;
      } else return -1;
    }
    return what == XG_OUT_TANGENT ? 36 : (what == XG_OUT_PLASTIC ? 7 : 6);
  }
};

struct OQuad : OElement {
  J2Params mp;
  double thk, b[2], rho;
  J2Point pts[4];
  int ndof() const override { return 8; }
  void gather(Dom& d, double crd[4][2], double u[4][2]) {
    for (int i = 0; i < 4; i++) {
      const NodeRec& n = d.nodes[nodes[i]];
      for (int j = 0; j < 2; j++) {
        crd[i][j] = n.crd[j];
        u[i][j] = n.U[j];
      }
    }
  }
  int update(Dom& d) override {
    double crd[4][2], u[4][2];
    gather(d, crd, u);
    quad_update(mp, d.dt, pts, crd, u);
    return 0;
  }
  void tangent(Dom& d, double* K) override {
    double crd[4][2], u[4][2];
    gather(d, crd, u);
    quad_tangent(pts, crd, thk, K);
  }
  void resist(Dom& d, double* R) override {
    double crd[4][2], u[4][2];
    gather(d, crd, u);
    quad_resisting_force(pts, crd, thk, b, R);
  }
  void commit() override {
    for (auto& p : pts) j2_commit(p);
  }
  void revert() override {}
  void revert_to_start() override {
    for (auto& p : pts) p.zero();
  }
  int npoints() const override { return 4; }
  int point_data(int what, double* out) const override {
    for (int g = 0; g < 4; g++) {
      if (what == XG_OUT_STRESS) j2_ps_get_stress(pts[g], out + 3 * g);
      else if (what == XG_OUT_TANGENT) j2_ps_get_tangent(pts[g], out + 9 * g);
      else if (what == XG_OUT_STRAIN) {
        const J2Point& s = pts[g];
        double* o = out + 3 * g;
        o[0] = s.strain[0][0]; o[1] = s.strain[1][1]; o[2] = 2.0 * s.strain[0][1];
      } else if (what == XG_OUT_PLASTIC) {
        const J2Point& s = pts[g];
        double* o = out + 7 * g;
        o[0] = s.epsp_np1[0][0]; o[1] = s.epsp_np1[1][1]; o[2] = s.epsp_np1[2][2];
        o[3] = s.epsp_np1[0][1]; o[4] = s.epsp_np1[1][2]; o[5] = s.epsp_np1[2][0];
        o[6] = s.xi_np1;
      } else return -1;
    }
    return what == XG_OUT_TANGENT ? 9 : (what == XG_OUT_PLASTIC ? 7 : 3);
  }
};

struct OBeam : OElement {
  FrameElement fe;
  int ndof() const override { return fe.ndof(); }
  void gather(Dom& d, double crdI[3], double crdJ[3], double ug[12]) {
    const NodeRec& a = d.nodes[nodes[0]];
    const NodeRec& c = d.nodes[nodes[1]];
    int nd = fe.is3d ? 6 : 3;
    for (int j = 0; j < 3; j++) {
      crdI[j] = a.crd[j];
      crdJ[j] = c.crd[j];
    }
    for (int j = 0; j < nd; j++) {
      ug[j] = a.U[j];
      ug[nd + j] = c.U[j];
    }
  }
  int update(Dom& d) override {
    double ug[12];
    double ci[3], cj[3];
    gather(d, ci, cj, ug);
    return fe.update(ug);
  }
  void tangent(Dom&, double* K) override { fe.tangent(K); }
  void resist(Dom&, double* R) override { fe.resisting_force(R); }
  void mass(Dom&, double* M) override { fe.lumped_mass(M); }
  void commit() override { fe.commit(); }
  void revert() override { fe.revert(); }
  void revert_to_start() override { fe.revert_to_start(); }
  int npoints() const override { return fe.nsec; }
  int point_data(int what, double* out) const override { return fe.point_data(what, out); }
};

struct OForce : OElement {
  ForceElement fb;
  int ndof() const override { return fb.fe.ndof(); }
  int update(Dom& d) override {
    const NodeRec& a = d.nodes[nodes[0]];
    const NodeRec& c = d.nodes[nodes[1]];
    const int nd = fb.fe.is3d ? 6 : 3;
    double ug[12];
    for (int j = 0; j < nd; j++) {
      ug[j] = a.U[j];
      ug[nd + j] = c.U[j];
    }
    return fb.update(ug);
  }
  void tangent(Dom&, double* K) override { fb.tangent(K); }
  void resist(Dom&, double* R) override { fb.resisting_force(R); }
  void commit() override { fb.commit(); }
  void revert() override { fb.revert(); }
  void revert_to_start() override { fb.revert_to_start(); }
  int npoints() const override { return fb.fe.nsec; }
  int point_data(int what, double* out) const override { return fb.fe.point_data(what, out); }
};

// ---------------------------------------------------------------------------
// Numbering: PlainHandler.cpp:60-247, AnalysisModel::getDOFGroupGraph
// (AnalysisModel.cpp:355-398), RCM::number (RCM.cpp:67-259, GPS=false),
// DOF_Numberer::numberDOF (DOF_Numberer.cpp:93-205), getDOFGraph
// (AnalysisModel.cpp:286-352) and SparseGenRowLinSOE::setSize (:129-234).
// ---------------------------------------------------------------------------
// sorted-unique adjacency insert, ID::insert (ID.cpp:512-560)
static void id_insert(std::vector<int>& v, int x) {
  auto it = std::lower_bound(v.begin(), v.end(), x);
  if (it != v.end() && *it == x) return;
  v.insert(it, x);
}

static std::vector<int> rcm_number(const std::vector<std::vector<int>>& adj) {
  int numVertex = (int)adj.size();
  std::vector<int> result(numVertex), tmp(numVertex, -1);
  if (numVertex == 0) return result;
  int iter4 = 0;  // VertexIter over ascending vertex tags
  int start = 0;  // first vertex from the iterator
  int currentMark = numVertex - 1;
  int nextMark = currentMark - 1;
  result[currentMark] = start;
  tmp[start] = currentMark;
  while (nextMark >= 0) {
    int v = result[currentMark];
    for (int w : adj[v]) {
      if (tmp[w] == -1) {
        tmp[w] = nextMark;
        result[nextMark--] = w;
      }
    }
    currentMark--;
    if ((currentMark == nextMark) && (currentMark >= 0)) {
      // loop over iter till we get a vertex not yet Tmped (RCM.cpp:225-233)
      int vp = -1;
      while (iter4 < numVertex) {
        int cand = iter4++;
        if (tmp[cand] == -1) {
          vp = cand;
          break;
        }
      }
      nextMark--;
      tmp[vp] = currentMark;
      result[currentMark] = vp;
    }
  }
  return result;  // result[i] = vertex tag at position i
}

// Element::setRayleighDampingFactors (Element.cpp:110-178): with betaKcomm
// != 0 an element without a committed tangent takes its current tangent,
// without it the stored one is dropped.
static void rayleigh_kc_init(Dom& d) {
  for (auto* e : d.eles) {
    if (d.betaKcomm == 0.0) {
      e->Kc.clear();
    } else if (e->Kc.empty()) {
      e->Kc.assign((size_t)e->ndof() * e->ndof(), 0.0);
      e->tangent(d, e->Kc.data());
    }
  }
}

static int setup(Dom& d) {
  if (d.nodemap.empty()) return fail(XG_ERR_STATE, "no nodes");
  d.nodes.clear();
  d.node_index.clear();
  for (auto& kv : d.nodemap) {
    d.node_index[kv.first] = (int)d.nodes.size();
    d.nodes.push_back(kv.second);
  }
  d.eles.clear();
  for (auto& kv : d.elemap) d.eles.push_back(kv.second.get());
  // resolve element node tags to indices
  for (auto* e : d.eles)
    for (auto& n : e->nodes) {
      auto it = d.node_index.find(n);
      if (it == d.node_index.end()) return fail(XG_ERR_ARG, "element node not found");
      n = it->second;
    }
  int N = (int)d.nodes.size();
  // DOF groups: -2 free, -1 fixed
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) n.eq[j] = n.fix[j] ? -1 : -2;
  // ordering of DOF groups
  std::vector<int> ordered(N);
  if (d.numberer == XG_NUMBERER_RCM) {
    std::vector<std::vector<int>> adj(N);
    for (auto* e : d.eles) {
      int sz = (int)e->nodes.size();
      for (int i = 0; i < sz; i++)
        for (int j = 0; j < sz; j++)
          if (i != j) {
            int a = e->nodes[i], b = e->nodes[j];
            if (a != b) {
              id_insert(adj[a], b);
              id_insert(adj[b], a);
            }
          }
    }
    ordered = rcm_number(adj);
  } else {
    for (int i = 0; i < N; i++) ordered[i] = i;  // PlainNumberer: node order
  }
  int eqn = 0;
  for (int i = 0; i < N; i++) {
    NodeRec& n = d.nodes[ordered[i]];
    for (int j = 0; j < d.ndf; j++)
      if (n.eq[j] == -2) n.eq[j] = eqn++;
  }
  d.neq = eqn;
  // DOF graph (vertex = equation number), sorted adjacency
  std::vector<std::vector<int>> dadj(d.neq);
  std::vector<int> id;
  for (auto* e : d.eles) {
    id.clear();
    for (int n : e->nodes)
      for (int j = 0; j < d.ndf; j++) id.push_back(d.nodes[n].eq[j]);
    int sz = (int)id.size();
    for (int i = 0; i < sz; i++) {
      int e1 = id[i];
      if (e1 < 0) continue;
      for (int j = i + 1; j < sz; j++) {
        int e2 = id[j];
        if (e2 < 0 || e2 == e1) continue;
        id_insert(dadj[e1], e2);
        id_insert(dadj[e2], e1);
      }
    }
  }
  d.rowptr.assign(d.neq + 1, 0);
  d.colind.clear();
  for (int a = 0; a < d.neq; a++) {
    std::vector<int> row = dadj[a];
    id_insert(row, a);
    d.colind.insert(d.colind.end(), row.begin(), row.end());
    d.rowptr[a + 1] = (int)d.colind.size();
  }
  d.Aval.assign(d.colind.size(), 0.0);
  d.B.assign(d.neq, 0.0);
  d.X.assign(d.neq, 0.0);
  // every element's setDomain: frame elements call update() (DispBeamColumn2d.cpp:184)
  for (auto* e : d.eles)
    if (dynamic_cast<OBeam*>(e)) e->update(d);
  d.ready = true;
  rayleigh_kc_init(d);
  return 0;
}

static void element_id(const Dom& d, const OElement* e, std::vector<int>& id) {
  id.clear();
  for (int n : e->nodes)
    for (int j = 0; j < d.ndf; j++) id.push_back(d.nodes[n].eq[j]);
}

// position of (row, col) in the CSR pattern
static long csr_pos(const Dom& d, int row, int col) {
  auto b = d.colind.begin() + d.rowptr[row], en = d.colind.begin() + d.rowptr[row + 1];
  auto it = std::lower_bound(b, en, col);
  return it - d.colind.begin();
}

// ---------------------------------------------------------------------------
// Domain operations
// ---------------------------------------------------------------------------
static int domain_update(Dom& d) {
  int res = 0;
  // Domain::update is a serial element loop (Domain.cpp:1988-2006); the
  // elements are independent, so a build with OpenMP (the all-cores CPU
  // baseline) splits the loop over threads with the same per-element results
  const long ne = (long)d.eles.size();
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : res) schedule(static)
#endif
  for (long i = 0; i < ne; i++) res += d.eles[i]->update(d);
  d.counters.updates++;
  for (auto* e : d.eles) d.counters.material_points += e->npoints();
  return res;
}

// Newmark/transient coefficients for formEleTangent (Newmark.cpp:262-289)
struct TangentFactors {
  double cK, cM, cC;
};

static TangentFactors tangent_factors(const Dom& d) {
  if (d.integ == 3) {
    double c1 = 1.0, c2 = d.nm_gamma / (d.nm_beta * d.dt), c3 = 1.0 / (d.nm_beta * d.dt * d.dt);
    return {c1, c3, c2};
  }
  return {1.0, 0.0, 0.0};
}

static int form_tangent(Dom& d) {
  std::fill(d.Aval.begin(), d.Aval.end(), 0.0);
  TangentFactors f = tangent_factors(d);
  std::vector<int> id;
  std::vector<double> K, M;
  for (auto* e : d.eles) {
    int nd = e->ndof();
    K.assign(nd * nd, 0.0);
    e->tangent(d, K.data());
    if (d.integ == 3) {
      // Newmark::formEleTangent (Newmark.cpp:262-289): T = c1 K, T += c2 C,
      // T += c3 M, with C = Element::getDamp (Element.cpp:181-200):
      // alphaM M, then + betaK K, + betaKcomm Kc
      M.assign(nd * nd, 0.0);
      e->mass(d, M.data());
      for (int i = 0; i < nd * nd; i++) {
        double c = 0.0;
        if (d.alphaM != 0.0) c = d.alphaM * M[i];
        if (d.betaK != 0.0) c = c + d.betaK * K[i];
        if (d.betaKcomm != 0.0) c = c + d.betaKcomm * e->Kc[i];
        double t = f.cK * K[i];
        t = t + f.cC * c;
        K[i] = t + f.cM * M[i];
      }
    }
    element_id(d, e, id);
    for (int i = 0; i < nd; i++) {
      int row = id[i];
      if (row < 0) continue;
      for (int j = 0; j < nd; j++) {
        int col = id[j];
        if (col < 0) continue;
        d.Aval[csr_pos(d, row, col)] += K[i * nd + j];
      }
    }
  }
  // nodal mass contributions (DOF_Group::addMtoTang), Newmark only
  if (d.integ == 3) {
    for (auto& n : d.nodes)
      for (int j = 0; j < d.ndf; j++)
        if (n.eq[j] >= 0 && n.mass[j] != 0.0) {
          int a = n.eq[j];
          d.Aval[csr_pos(d, a, a)] += (f.cM + f.cC * d.alphaM) * n.mass[j];
        }
  }
  return 0;
}

// ground acceleration by PathSeries linear interpolation (PathSeries.cpp:334-356)
static double ground_accel(const Dom& d, double t) {
  if (d.gm_accel.empty() || d.gm_dt <= 0) return 0.0;
  double incr = t / d.gm_dt;
  int i1 = (int)incr;
  int i2 = i1 + 1;
  int n = (int)d.gm_accel.size();
  if (i2 >= n) return 0.0;
  double v1 = d.gm_accel[i1], v2 = d.gm_accel[i2];
  return d.gm_factor * (v1 + (v2 - v1) * (incr - i1));
}

static int form_unbalance(Dom& d) {
  std::fill(d.B.begin(), d.B.end(), 0.0);
  std::vector<int> id;
  std::vector<double> R, M;
  for (auto* e : d.eles) {
    int nd = e->ndof();
    R.assign(nd, 0.0);
    e->resist(d, R.data());
    if (d.integ == 3) {
      // getResistingForceIncInertia: + M a (lumped), then + the Rayleigh
      // damping force D v (Element::getRayleighDampingForces,
      // Element.cpp:283-323: D = alphaM M, + betaK K, + betaKcomm Kc)
      M.assign(nd * nd, 0.0);
      e->mass(d, M.data());
      int nn = (int)e->nodes.size();
      std::vector<double> v(nd);
      for (int a = 0; a < nn; a++)
        for (int j = 0; j < d.ndf; j++) {
          const NodeRec& nr = d.nodes[e->nodes[a]];
          int i = a * d.ndf + j;
          R[i] += M[i * nd + i] * nr.A[j];
          v[i] = nr.V[j];
        }
      if (d.alphaM != 0.0 || d.betaK != 0.0 || d.betaKcomm != 0.0) {
        std::vector<double> Kt;
        if (d.betaK != 0.0) {
          Kt.assign((size_t)nd * nd, 0.0);
          e->tangent(d, Kt.data());
        }
        for (int i = 0; i < nd; i++) {
          double fi = 0.0;
          for (int j = 0; j < nd; j++) {
            double c = 0.0;
            if (d.alphaM != 0.0) c = d.alphaM * M[i * nd + j];
            if (d.betaK != 0.0) c = c + d.betaK * Kt[i * nd + j];
            if (d.betaKcomm != 0.0) c = c + d.betaKcomm * e->Kc[i * nd + j];
            fi += c * v[j];
          }
          R[i] += fi;
        }
      }
    }
    element_id(d, e, id);
    for (int i = 0; i < nd; i++)
      if (id[i] >= 0) d.B[id[i]] += -R[i];
  }
  // nodal unbalance in DOF-group order (IncrementalIntegrator.cpp:202-218)
  double ag = (d.integ == 3) ? ground_accel(d, d.time) : 0.0;
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) {
      int a = n.eq[j];
      if (a < 0) continue;
      double u = d.lambda * n.P[j];
      if (d.integ == 3 && n.mass[j] != 0.0) {
        // Node::addInertiaLoadToUnbalance (-M R ag, Node.cpp:668-688), then
        // Node::getUnbalancedLoadIncInertia (-M A, -alphaM M V; :740-763)
        const double r = (j == d.gm_dir) ? 1.0 : 0.0;
        if (r != 0.0) u += -(n.mass[j] * r) * ag;
        u += -n.mass[j] * n.A[j];
        if (d.alphaM != 0.0) u += -d.alphaM * n.mass[j] * n.V[j];
      }
      d.B[a] += u;
    }
  if (d.integ == 3 && d.gm_dir >= 0) {
    // -M_e r ag for element (lumped) masses
    for (auto* e : d.eles) {
      int nd = e->ndof();
      M.assign(nd * nd, 0.0);
      e->mass(d, M.data());
      element_id(d, e, id);
      int nn = (int)e->nodes.size();
      for (int a = 0; a < nn; a++) {
        int i = a * d.ndf + d.gm_dir;
        if (id[i] >= 0) d.B[id[i]] -= M[i * nd + i] * ag;
      }
    }
  }
  return 0;
}

// Profile LDL^T of the assembled CSR matrix (lower profile by rows).  This
// plays the role of the reference's default ProfileSPDLinSOE direct solve.
static int direct_solve(Dom& d, const std::vector<double>& b, std::vector<double>& x) {
  int n = d.neq;
  std::vector<int> first(n);
  for (int i = 0; i < n; i++) first[i] = d.colind[d.rowptr[i]];
  std::vector<long> start(n + 1);
  start[0] = 0;
  for (int i = 0; i < n; i++) start[i + 1] = start[i] + (i - first[i] + 1);
  std::vector<double> L(start[n], 0.0);
  for (int i = 0; i < n; i++)
    for (int k = d.rowptr[i]; k < d.rowptr[i + 1]; k++) {
      int j = d.colind[k];
      if (j <= i) L[start[i] + (j - first[i])] = d.Aval[k];
    }
  auto at = [&](int i, int j) -> double& { return L[start[i] + (j - first[i])]; };
  std::vector<double> D(n);
  for (int i = 0; i < n; i++) {
    for (int j = first[i]; j < i; j++) {
      int kmin = std::max(first[i], first[j]);
      double s = at(i, j);
      for (int k = kmin; k < j; k++) s -= at(i, k) * at(j, k) * D[k];
      at(i, j) = s / D[j];
    }
    double s = at(i, i);
    for (int k = first[i]; k < i; k++) s -= at(i, k) * at(i, k) * D[k];
    if (s == 0.0 || !std::isfinite(s)) return fail(XG_BAD_LINEAR_SOLVE, "singular matrix in direct solve");
    D[i] = s;
  }
  x = b;
  for (int i = 0; i < n; i++) {
    double s = x[i];
    for (int k = first[i]; k < i; k++) s -= at(i, k) * x[k];
    x[i] = s;
  }
  for (int i = 0; i < n; i++) x[i] /= D[i];
  for (int i = n - 1; i >= 0; i--) {
    double xi = x[i];
    for (int k = first[i]; k < i; k++) x[k] -= at(i, k) * xi;
  }
  d.counters.solves++;
  return 0;
}

static int solve(Dom& d) {
  int r = direct_solve(d, d.B, d.X);
  return r;
}

static void incr_disp(Dom& d, const std::vector<double>& dU) {
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++)
      if (n.eq[j] >= 0) n.U[j] += dU[n.eq[j]];
}

static void commit_domain(Dom& d) {
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) {
      n.Uc[j] = n.U[j];
      n.Vc[j] = n.V[j];
      n.Ac[j] = n.A[j];
    }
  for (auto* e : d.eles) {
    // Element::commitState (Element.cpp:74-80), called first by each
    // element's own commitState: Kc = getTangentStiff()
    if (!e->Kc.empty()) e->tangent(d, e->Kc.data());
    e->commit();
  }
}

// Domain::revertToLastCommit (Domain.cpp:1925-1939): nodes and elements
// revert, then the domain updates once more at the committed displacements
// with dT = 0.
static void revert_domain(Dom& d) {
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) {
      n.U[j] = n.Uc[j];
      n.V[j] = n.Vc[j];
      n.A[j] = n.Ac[j];
    }
  for (auto* e : d.eles) e->revert();
  const double dt_keep = d.dt;
  d.dt = 0.0;
  domain_update(d);
  d.dt = dt_keep;
}

// ---- convergence tests ------------------------------------------------------
struct Test {
  int currentIter = 0, numIncr = 0;
  std::vector<double> norms;
};

static double norm2(const std::vector<double>& v) {
  double s = 0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

// returns >0 converged (iteration count), -1 continue, -2 failure
static int run_test(Dom& d, Test& t) {
  double norm;
  if (d.test_kind == XG_TEST_NORM_UNBALANCE) norm = norm2(d.B);
  else if (d.test_kind == XG_TEST_NORM_DISP_INCR) norm = norm2(d.X);
  else {
    double s = 0;
    for (int i = 0; i < d.neq; i++) s += d.X[i] * d.B[i];
    norm = 0.5 * std::fabs(s);
  }
  d.counters.last_norm = norm;
  if (t.currentIter > 1 && !t.norms.empty() && t.norms.back() < norm) t.numIncr++;
  t.norms.push_back(norm);
  if (norm <= d.test_tol) return t.currentIter;
  if (t.currentIter >= d.test_maxit || (d.test_kind == XG_TEST_NORM_UNBALANCE && t.numIncr >= d.test_maxit)) {
    t.currentIter++;
    return -2;
  }
  t.currentIter++;
  return -1;
}

// ---- integrators --------------------------------------------------------------
static int integ_update(Dom& d, const std::vector<double>& dUin);

static int newmark_update_response(Dom& d, const std::vector<double>& deltaU);

static int integ_new_step(Dom& d, double dt) {
  if (d.integ == 1) {  // LoadControl::newStep, LoadControl.cpp:70-95
    double factor = d.lc_spec / d.lc_last;
    d.lc_dlambda *= factor;
    if (d.lc_dlambda < d.lc_dlmin) d.lc_dlambda = d.lc_dlmin;
    else if (d.lc_dlambda > d.lc_dlmax) d.lc_dlambda = d.lc_dlmax;
    d.lambda += d.lc_dlambda;
    d.lc_last = 0;
    return 0;
  }
  if (d.integ == 2) {  // DisplacementControl::newStep, DisplacementControl.cpp:121-222
    double factor = d.dc_spec / d.dc_last;
    d.dc_incr *= factor;
    if (d.dc_incr < d.dc_min) d.dc_incr = d.dc_min;
    else if (d.dc_incr > d.dc_max) d.dc_incr = d.dc_max;
    form_tangent(d);
    if (direct_solve(d, d.phat, d.dUhat) < 0) return XG_BAD_STEP;
    double dUahat = d.dUhat[d.dc_dofid];
    if (dUahat == 0.0) return fail(XG_BAD_STEP, "dUahat is zero");
    double dlambda = d.dc_incr / dUahat;
    d.dc_dlambda_step = dlambda;
    d.lambda += dlambda;
    for (int i = 0; i < d.neq; i++) d.dU[i] = d.dUhat[i] * dlambda;
    d.dUstep = d.dU;
    incr_disp(d, d.dU);
    if (domain_update(d) < 0) return XG_BAD_STEP;
    d.dc_last = 0;
    return 0;
  }
  // Newmark::newStep, Newmark.cpp:105-240 (displacement-free predictor)
  d.dt = dt;
  double c1 = 1.0, c2 = d.nm_gamma / (d.nm_beta * dt), c3 = 1.0 / (d.nm_beta * dt * dt);
  (void)c1; (void)c2; (void)c3;
  double a1 = (1.0 - d.nm_gamma / d.nm_beta);
  double a2 = dt * (1.0 - 0.5 * d.nm_gamma / d.nm_beta);
  double a3 = -1.0 / (d.nm_beta * dt);
  double a4 = 1.0 - 0.5 / d.nm_beta;
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) {
      double Vn = n.Vc[j], An = n.Ac[j];
      n.U[j] = n.Uc[j];
      n.V[j] = a1 * Vn + a2 * An;
      n.A[j] = a3 * Vn + a4 * An;
    }
  d.time += dt;
  return domain_update(d) < 0 ? XG_BAD_STEP : 0;
}

static int integ_update(Dom& d, const std::vector<double>& dUin) {
  if (d.integ == 1) {  // LoadControl::update, LoadControl.cpp:102-120
    incr_disp(d, dUin);
    if (domain_update(d) < 0) return XG_BAD_STEP;
    d.lc_last++;
    return 0;
  }
  if (d.integ == 2) {  // DisplacementControl::update, DisplacementControl.cpp:224-283
    d.dUbar = dUin;
    double dUabar = d.dUbar[d.dc_dofid];
    if (direct_solve(d, d.phat, d.dUhat) < 0) return XG_BAD_STEP;
    double dUahat = d.dUhat[d.dc_dofid];
    if (dUahat == 0.0) return fail(XG_BAD_STEP, "dUahat is zero");
    double dLambda = -dUabar / dUahat;
    for (int i = 0; i < d.neq; i++) d.dU[i] = d.dUbar[i] + dLambda * d.dUhat[i];
    for (int i = 0; i < d.neq; i++) d.dUstep[i] += d.dU[i];
    d.dc_dlambda_step += dLambda;
    d.lambda += dLambda;
    incr_disp(d, d.dU);
    if (domain_update(d) < 0) return XG_BAD_STEP;
    d.X = d.dU;
    d.dc_last++;
    return 0;
  }
  return newmark_update_response(d, dUin);
}

// Newmark::update, Newmark.cpp:411-458
static int newmark_update_response(Dom& d, const std::vector<double>& deltaU) {
  double c2 = d.nm_gamma / (d.nm_beta * d.dt), c3 = 1.0 / (d.nm_beta * d.dt * d.dt);
  for (auto& n : d.nodes)
    for (int j = 0; j < d.ndf; j++) {
      int a = n.eq[j];
      if (a < 0) continue;
      n.U[j] += deltaU[a];
      n.V[j] += c2 * deltaU[a];
      n.A[j] += c3 * deltaU[a];
    }
  return domain_update(d) < 0 ? XG_BAD_STEP : 0;
}

// NewtonRaphson::solveCurrentStep, NewtonRaphson.cpp:125-221
static int newton(Dom& d) {
  if (form_unbalance(d) < 0) return XG_BAD_STEP;
  Test t;
  t.currentIter = 1;
  int result = -1;
  do {
    if (form_tangent(d) < 0) return XG_BAD_STEP;
    if (solve(d) < 0) return XG_BAD_LINEAR_SOLVE;
    if (integ_update(d, d.X) < 0) return XG_BAD_STEP;
    if (form_unbalance(d) < 0) return XG_BAD_STEP;
    result = run_test(d, t);
    d.counters.newton_iters++;
  } while (result == -1);
  if (result == -2) return fail(XG_TEST_FAILED, "convergence test failed");
  return 0;
}

static int analyze(Dom& d, int nsteps, double dt) {
  if (!d.ready) {
    int r = setup(d);
    if (r < 0) return r;
  }
  if (d.integ == 2 && d.phat.empty()) {
    // DisplacementControl::domainChanged, :286-415
    d.phat.assign(d.neq, 0.0);
    d.dUhat.assign(d.neq, 0.0);
    d.dUbar.assign(d.neq, 0.0);
    d.dU.assign(d.neq, 0.0);
    d.dUstep.assign(d.neq, 0.0);
    d.lambda += 1.0;
    form_unbalance(d);
    d.phat = d.B;
    d.lambda -= 1.0;
    // check there is a reference load (DisplacementControl.cpp:386-398)
    bool have_load = false;
    for (double v : d.phat)
      if (v != 0.0) {
        have_load = true;
        break;
      }
    if (!have_load) {
      d.phat.clear();
      return fail(XG_ERR_ARG, "DisplacementControl: zero reference load");
    }
    auto it = d.node_index.find(d.dc_node);
    if (it == d.node_index.end()) return fail(XG_ERR_ARG, "control node not found");
    d.dc_dofid = d.nodes[it->second].eq[d.dc_dof];
    if (d.dc_dofid < 0) return fail(XG_ERR_ARG, "control dof is fixed");
  }
  for (int s = 0; s < nsteps; s++) {
    double lam0 = d.lambda, t0 = d.time;
    int r = integ_new_step(d, dt);
    if (r == 0) r = newton(d);
    if (r < 0) {
      revert_domain(d);
      d.lambda = lam0;
      d.time = t0;
      return r;
    }
    commit_domain(d);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Direct differentiation of a Newmark transient of a frame model (the
// forward-mode counterpart of the GPU adjoint).  Per converged step, as
// Newmark::computeSensitivities (Newmark.cpp:904-950): the sensitivity rhs
// (formSensitivityRHS, :691) = -(dF_int/dtheta at fixed displacement, with
// the history sensitivities of the last commit) - M (dA_pred) - alphaM M
// (dV_pred), one solve with the converged effective tangent, the response
// sensitivities (saveSensitivity, :791) and the material history
// sensitivities at the converged strain (commitSensitivity, :880).  The
// material partials, absent from the reference, come from the forward-mode
// restatement in orc_ddm.hpp.
// ---------------------------------------------------------------------------
struct DdmDir {
  std::map<int, SteelTheta<double>> steel;  // material tag -> (dFy, dE0, db)
  std::map<int, double> conc;               // material tag -> dfc
};

struct DdmState {
  std::vector<double> hdot;  // per fiber, 10 history sensitivities
  std::vector<double> Ud, Vd, Ad;  // per node x 6
};

static double section_strain(const FrameElement& fe, const FiberSection& S, const FiberDef& F, const double* dfm) {
  const double y = F.y - S.yBar;
  if (!fe.is3d) return dfm[0] - y * dfm[1];
  const double z = F.z - S.zBar;
  return dfm[0] - y * dfm[1] + z * dfm[2];
}

static void section_deformations(const FrameElement& fe, const double* ug, double dfm[5][4]) {
  double v[6];
  fe.basic_disp(ug, v);
  const double oneOverL = 1.0 / fe.L;
  for (int i = 0; i < fe.nsec; i++) {
    const double xi6 = 6.0 * (0.5 * (1.0 + kGL5pts[i]));
    dfm[i][0] = oneOverL * v[0];
    dfm[i][1] = oneOverL * ((xi6 - 4.0) * v[1] + (xi6 - 2.0) * v[2]);
    dfm[i][2] = dfm[i][3] = 0.0;
    if (fe.is3d) {
      dfm[i][2] = oneOverL * ((xi6 - 4.0) * v[3] + (xi6 - 2.0) * v[4]);
      dfm[i][3] = oneOverL * v[5];
    }
  }
}

// one fiber: committed state (+ history sensitivities) and strain (+ its
// sensitivity) -> trial state sensitivities in pool order; returns dsigma
static double fiber_ddm(const DdmDir& dir, const FiberDef& F, const FiberState& st, double strain, double dstrain,
                        const double* hd, double* hout) {
  if (F.kind == 1) {
    SteelTheta<Dual> th{Dual(F.steel.Fy), Dual(F.steel.E0), Dual(F.steel.b)};
    auto it = dir.steel.find(F.mat_tag);
    if (it != dir.steel.end()) {
      th.Fy.d = it->second.Fy;
      th.E0.d = it->second.E0;
      th.b.d = it->second.b;
    }
    const Steel02State& s = st.st;
    SteelH<Dual> c{{s.epsminP, hd[0]}, {s.epsmaxP, hd[1]}, {s.epsplP, hd[2]}, {s.epss0P, hd[3]},
                   {s.sigs0P, hd[4]},  {s.epssrP, hd[5]}, {s.sigsrP, hd[6]}, {s.epsP, hd[7]},
                   {s.sigP, hd[8]},    {s.eP, hd[9]},     s.konP};
    SteelH<Dual> t;
    steel02_trial_t(F.steel, th, c, Dual(strain, dstrain), t);
    const Dual o[10] = {t.epsmin, t.epsmax, t.epspl, t.epss0, t.sigs0, t.epsr, t.sigr, t.eps, t.sig, t.e};
    for (int k = 0; k < 10; k++) hout[k] = o[k].d;
    return t.sig.d;
  }
  if (F.kind == 2) {
    Dual fc(F.conc.fc);
    auto it = dir.conc.find(F.mat_tag);
    if (it != dir.conc.end()) fc.d = it->second;
    const Concrete02State& s = st.cs;
    ConcH<Dual> c{{s.ecminP, hd[0]}, {s.deptP, hd[1]}, {s.epsP, hd[2]}, {s.sigP, hd[3]}, {s.eP, hd[4]}};
    ConcH<Dual> t;
    concrete02_trial_t(F.conc, fc, c, c.sig, c.e, Dual(strain, dstrain), t);
    const Dual o[5] = {t.ecmin, t.dept, t.eps, t.sig, t.e};
    for (int k = 0; k < 5; k++) hout[k] = o[k].d;
    for (int k = 5; k < 10; k++) hout[k] = 0.0;
    return t.sig.d;
  }
  for (int k = 0; k < 10; k++) hout[k] = 0.0;
  return F.E * dstrain;
}

static int ddm_step(Dom& d, const DdmDir& dir, DdmState& S) {
  const double g = d.nm_gamma, b = d.nm_beta, h = d.dt;
  const double c2 = g / (b * h), c3 = 1.0 / (b * h * h);
  const double a1 = 1.0 - g / b, a2 = h * (1.0 - 0.5 * g / b), a3 = -1.0 / (b * h), a4 = 1.0 - 0.5 / b;
  const int ndf = d.ndf;
  std::vector<double> rhs(d.neq, 0.0), x;
  std::vector<int> id;
  double hout[10];
  // -dF_int/dtheta at the converged displacement
  size_t fib = 0;
  for (auto* e : d.eles) {
    auto* be = dynamic_cast<OBeam*>(e);
    if (!be) return fail(XG_ERR_ARG, "ddm: frame elements only");
    FrameElement& fe = be->fe;
    if (fe.rho != 0.0) return fail(XG_ERR_ARG, "ddm: element mass density is not supported");
    double ci[3], cj[3], ug[12];
    be->gather(d, ci, cj, ug);
    double dfm[5][4];
    section_deformations(fe, ug, dfm);
    double saved[5][4];
    for (int i = 0; i < fe.nsec; i++) {
      FiberSection& sec = fe.secs[i];
      double sd[4] = {0, 0, 0, 0};
      for (size_t f = 0; f < sec.fs.size(); f++, fib++) {
        const FiberDef& F = sec.def->fibers[f];
        const double eps = section_strain(fe, sec, F, dfm[i]);
        const double dsig = fiber_ddm(dir, F, sec.fs[f], eps, 0.0, &S.hdot[fib * 10], hout);
        const double y = F.y - sec.yBar, z = fe.is3d ? F.z - sec.zBar : 0.0;
        sd[0] += F.A * dsig;
        sd[1] += -y * F.A * dsig;
        sd[2] += z * F.A * dsig;
      }
      for (int k = 0; k < 4; k++) {
        saved[i][k] = sec.s[k];
        sec.s[k] = fe.is3d ? sd[k] : (k < 2 ? sd[k] : 0.0);
      }
    }
    std::vector<double> P(fe.ndof());
    fe.resisting_force(P.data());
    for (int i = 0; i < fe.nsec; i++)
      for (int k = 0; k < 4; k++) fe.secs[i].s[k] = saved[i][k];
    element_id(d, e, id);
    for (int i = 0; i < fe.ndof(); i++)
      if (id[i] >= 0) rhs[id[i]] -= P[i];
  }
  // inertia and damping terms of the previous step's sensitivities
  for (size_t k = 0; k < d.nodes.size(); k++) {
    const NodeRec& n = d.nodes[k];
    for (int j = 0; j < ndf; j++) {
      const int a = n.eq[j];
      if (a < 0 || n.mass[j] == 0.0) continue;
      const double Ud = S.Ud[k * 6 + j], Vd = S.Vd[k * 6 + j], Ad = S.Ad[k * 6 + j];
      const double apred = -c3 * Ud + a3 * Vd + a4 * Ad;
      const double vpred = -c2 * Ud + a1 * Vd + a2 * Ad;
      rhs[a] -= n.mass[j] * apred + d.alphaM * n.mass[j] * vpred;
    }
  }
  if (form_tangent(d) < 0) return XG_BAD_STEP;
  if (direct_solve(d, rhs, x) < 0) return XG_BAD_LINEAR_SOLVE;
  for (size_t k = 0; k < d.nodes.size(); k++) {
    const NodeRec& n = d.nodes[k];
    for (int j = 0; j < ndf; j++) {
      const int a = n.eq[j];
      const double Un = a >= 0 ? x[a] : 0.0;
      const double Ud = S.Ud[k * 6 + j], Vd = S.Vd[k * 6 + j], Ad = S.Ad[k * 6 + j];
      S.Vd[k * 6 + j] = c2 * (Un - Ud) + a1 * Vd + a2 * Ad;
      S.Ad[k * 6 + j] = c3 * (Un - Ud) + a3 * Vd + a4 * Ad;
      S.Ud[k * 6 + j] = Un;
    }
  }
  // history sensitivities at the converged strain (commitSensitivity)
  fib = 0;
  for (auto* e : d.eles) {
    auto* be = static_cast<OBeam*>(e);
    FrameElement& fe = be->fe;
    double ci[3], cj[3], ug[12], ugd[12];
    be->gather(d, ci, cj, ug);
    const int nd = fe.is3d ? 6 : 3;
    for (int a = 0; a < 2; a++)
      for (int j = 0; j < nd; j++) ugd[a * nd + j] = S.Ud[(size_t)e->nodes[a] * 6 + j];
    double dfm[5][4], dfmd[5][4];
    section_deformations(fe, ug, dfm);
    section_deformations(fe, ugd, dfmd);
    for (int i = 0; i < fe.nsec; i++) {
      FiberSection& sec = fe.secs[i];
      for (size_t f = 0; f < sec.fs.size(); f++, fib++) {
        const FiberDef& F = sec.def->fibers[f];
        const double eps = section_strain(fe, sec, F, dfm[i]);
        const double deps = section_strain(fe, sec, F, dfmd[i]);
        fiber_ddm(dir, F, sec.fs[f], eps, deps, &S.hdot[fib * 10], hout);
        for (int k = 0; k < 10; k++) S.hdot[fib * 10 + k] = hout[k];
      }
    }
  }
  return 0;
}

// loss = sum over steps of sum_i U(node_i, dof)^2 and its derivative along dir
static int ddm_transient(Dom& d, int nsteps, double dt, const DdmDir& dir, const std::vector<int>& loss_nodes,
                         int loss_dof, double& loss, double& dloss) {
  if (!d.ready) {
    int r = setup(d);
    if (r < 0) return r;
  }
  if (d.integ != 3) return fail(XG_ERR_STATE, "ddm: the Newmark integrator is required");
  if (d.betaK != 0 || d.betaKinit != 0 || d.betaKcomm != 0)
    return fail(XG_ERR_ARG, "ddm: stiffness-proportional damping is not supported");
  DdmState S;
  size_t nfib = 0;
  for (auto* e : d.eles) {
    auto* be = dynamic_cast<OBeam*>(e);
    if (!be) return fail(XG_ERR_ARG, "ddm: frame elements only");
    for (auto& sec : be->fe.secs) nfib += sec.fs.size();
  }
  S.hdot.assign(nfib * 10, 0.0);
  S.Ud.assign(d.nodes.size() * 6, 0.0);
  S.Vd = S.Ud;
  S.Ad = S.Ud;
  loss = dloss = 0.0;
  for (int s = 0; s < nsteps; s++) {
    const double t0 = d.time;
    int r = integ_new_step(d, dt);
    if (r == 0) r = newton(d);
    if (r == 0) r = ddm_step(d, dir, S);
    if (r < 0) {
      revert_domain(d);
      d.time = t0;
      return r;
    }
    for (int k : loss_nodes) {
      const double u = d.nodes[k].U[loss_dof];
      loss += u * u;
      dloss += 2.0 * u * S.Ud[(size_t)k * 6 + loss_dof];
    }
    commit_domain(d);
  }
  return 0;
}

}  // namespace orc

using namespace orc;

// =============================================================================
// C-ABI (same signatures as include/xara_gpu.h, prefix orc_)
// =============================================================================
struct xg_domain : orc::Dom {};

#define DOM(d) static_cast<orc::Dom*>(d)

extern "C" {

const char* orc_last_error(void) { return orc::g_err.c_str(); }

int orc_create(int ndm, int ndf, int device, xg_domain** out) {
  (void)device;
  if (ndm < 1 || ndm > 3 || ndf < 1 || ndf > 6) return fail(XG_ERR_ARG, "bad ndm/ndf");
  xg_domain* d = new xg_domain();
  d->ndm = ndm;
  d->ndf = ndf;
  *out = d;
  return 0;
}
int orc_destroy(xg_domain* d) {
  delete d;
  return 0;
}
int orc_nodes(xg_domain* d, int n, const int* tags, const double* crds) {
  for (int i = 0; i < n; i++) {
    NodeRec r;
    r.tag = tags[i];
    for (int j = 0; j < d->ndm; j++) r.crd[j] = crds[i * d->ndm + j];
    d->nodemap[tags[i]] = r;
  }
  d->ready = false;
  return 0;
}
int orc_fix(xg_domain* d, int n, const int* tags, const int* flags) {
  for (int i = 0; i < n; i++) {
    auto it = d->nodemap.find(tags[i]);
    if (it == d->nodemap.end()) return fail(XG_ERR_ARG, "fix: node not found");
    for (int j = 0; j < d->ndf; j++)
      if (flags[i * d->ndf + j]) it->second.fix[j] = 1;
  }
  d->ready = false;
  return 0;
}
int orc_mass(xg_domain* d, int n, const int* tags, const double* m) {
  for (int i = 0; i < n; i++) {
    auto it = d->nodemap.find(tags[i]);
    if (it == d->nodemap.end()) return fail(XG_ERR_ARG, "mass: node not found");
    for (int j = 0; j < d->ndf; j++) it->second.mass[j] = m[i * d->ndf + j];
  }
  return 0;
}
int orc_nd_j2plasticity(xg_domain* d, int tag, double K, double G, double sig0, double sigInf,
                        double delta, double H, double eta) {
  d->j2[tag] = J2Params{K, G, sig0, sigInf, delta, H, eta};
  return 0;
}
int orc_uniaxial_steel02(xg_domain* d, int tag, double fy, double E0, double b, double R0, double cR1,
                         double cR2, double a1, double a2, double a3, double a4, double sigInit) {
  d->steel[tag] = Steel02Params{fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sigInit};
  return 0;
}
int orc_uniaxial_concrete02(xg_domain* d, int tag, double fc, double epsc0, double fcu, double epscu,
                            double rat, double ft, double Ets) {
  d->conc[tag] = concrete02_make(fc, epsc0, fcu, epscu, rat, ft, Ets);
  return 0;
}
int orc_uniaxial_elastic(xg_domain* d, int tag, double E) {
  d->elastic[tag] = E;
  return 0;
}
int orc_section_fiber(xg_domain* d, int tag, int nfib, const int* mat_tags, const double* y,
                      const double* z, const double* A, double GJ, int compute_centroid) {
  FiberSectionDef s;
  s.is3d = (d->ndm == 3);
  s.GJ = GJ;
  s.compute_centroid = compute_centroid != 0;
  for (int i = 0; i < nfib; i++) {
    FiberDef f;
    f.y = y[i];
    f.z = z ? z[i] : 0.0;
    f.A = A[i];
    int mt = mat_tags[i];
    f.mat_tag = mt;
    if (d->steel.count(mt)) {
      f.kind = 1;
      f.steel = d->steel[mt];
    } else if (d->conc.count(mt)) {
      f.kind = 2;
      f.conc = d->conc[mt];
    } else if (d->elastic.count(mt)) {
      f.kind = 3;
      f.E = d->elastic[mt];
    } else
      return fail(XG_ERR_ARG, "section: unknown uniaxial material");
    s.fibers.push_back(f);
  }
  d->sections[tag] = s;
  return 0;
}
int orc_geom_transf_linear(xg_domain* d, int tag, const double* vecxz) {
  std::vector<double> v(3, 0.0);
  if (vecxz && d->ndm == 3) v = {vecxz[0], vecxz[1], vecxz[2]};
  d->transfs[tag] = v;
  return 0;
}
int orc_elements_brick(xg_domain* d, int n, const int* tags, const int* conn, int mat_tag, double b1,
                       double b2, double b3) {
  if (!d->j2.count(mat_tag)) return fail(XG_ERR_ARG, "stdBrick: unknown nDMaterial");
  if (d->ndm != 3 || d->ndf != 3) return fail(XG_ERR_ARG, "stdBrick needs ndm=ndf=3");
  for (int i = 0; i < n; i++) {
    auto e = std::make_unique<OBrick>();
    e->tag = tags[i];
    e->nodes.assign(conn + 8 * i, conn + 8 * i + 8);
    e->mp = d->j2[mat_tag];
    e->b[0] = b1; e->b[1] = b2; e->b[2] = b3;
    for (auto& p : e->pts) {
      p.zero();
      // the constructor of J2Plasticity runs plastic_integrator once on zero strain
      j2_plastic_integrator(e->mp, 0.0, p);
    }
    d->elemap[tags[i]] = std::move(e);
  }
  d->ready = false;
  return 0;
}
int orc_elements_quad(xg_domain* d, int n, const int* tags, const int* conn, double thk, int mat_tag,
                      double pressure, double rho, double b1, double b2) {
  if (!d->j2.count(mat_tag)) return fail(XG_ERR_ARG, "quad: unknown nDMaterial");
  if (d->ndm != 2 || d->ndf != 2) return fail(XG_ERR_ARG, "quad needs ndm=ndf=2");
  if (pressure != 0.0) return fail(XG_ERR_ARG, "quad: surface pressure is not supported");
  for (int i = 0; i < n; i++) {
    auto e = std::make_unique<OQuad>();
    e->tag = tags[i];
    e->nodes.assign(conn + 4 * i, conn + 4 * i + 4);
    e->mp = d->j2[mat_tag];
    e->thk = thk;
    e->rho = rho;
    e->b[0] = b1; e->b[1] = b2;
    for (auto& p : e->pts) {
      p.zero();
      j2_plastic_integrator(e->mp, 0.0, p);
    }
    d->elemap[tags[i]] = std::move(e);
  }
  d->ready = false;
  return 0;
}
int orc_elements_dispbeam(xg_domain* d, int n, const int* tags, const int* conn, int transf_tag,
                          int sec_tag, int nip, double rho) {
  if (!d->sections.count(sec_tag)) return fail(XG_ERR_ARG, "dispBeamColumn: unknown section");
  if (!d->transfs.count(transf_tag)) return fail(XG_ERR_ARG, "dispBeamColumn: unknown transf");
  bool is3d = d->ndm == 3;
  if ((is3d && d->ndf != 6) || (!is3d && d->ndf != 3)) return fail(XG_ERR_ARG, "dispBeamColumn ndf");
  if (nip != 5) return fail(XG_ERR_ARG, "dispBeamColumn: only 5 Legendre points are supported");
  for (int i = 0; i < n; i++) {
    auto e = std::make_unique<OBeam>();
    e->tag = tags[i];
    e->nodes.assign(conn + 2 * i, conn + 2 * i + 2);
    auto ni = d->nodemap.find(conn[2 * i]);
    auto nj = d->nodemap.find(conn[2 * i + 1]);
    if (ni == d->nodemap.end() || nj == d->nodemap.end()) return fail(XG_ERR_ARG, "beam node missing");
    int r = e->fe.init(is3d, ni->second.crd, nj->second.crd, d->transfs[transf_tag].data(),
                       d->sections[sec_tag], nip, rho);
    if (r < 0) return fail(XG_ERR_ARG, "beam init failed");
    d->elemap[tags[i]] = std::move(e);
  }
  d->ready = false;
  return 0;
}
int orc_elements_forcebeam(xg_domain* d, int n, const int* tags, const int* conn, int transf_tag, int sec_tag,
                           int nip, int max_iter, double tol) {
  if (!d->sections.count(sec_tag)) return fail(XG_ERR_ARG, "forceBeamColumn: unknown section");
  if (!d->transfs.count(transf_tag)) return fail(XG_ERR_ARG, "forceBeamColumn: unknown transf");
  bool is3d = d->ndm == 3;
  if ((is3d && d->ndf != 6) || (!is3d && d->ndf != 3)) return fail(XG_ERR_ARG, "forceBeamColumn ndf");
  if (nip < 2 || nip > 10) return fail(XG_ERR_ARG, "forceBeamColumn: 2 to 10 Lobatto points");
  for (int i = 0; i < n; i++) {
    auto e = std::make_unique<OForce>();
    e->tag = tags[i];
    e->nodes.assign(conn + 2 * i, conn + 2 * i + 2);
    auto ni = d->nodemap.find(conn[2 * i]);
    auto nj = d->nodemap.find(conn[2 * i + 1]);
    if (ni == d->nodemap.end() || nj == d->nodemap.end()) return fail(XG_ERR_ARG, "beam node missing");
    int r = e->fb.init(is3d, ni->second.crd, nj->second.crd, d->transfs[transf_tag].data(), d->sections[sec_tag],
                       nip, max_iter, tol);
    if (r == -4) return fail(XG_ERR_ARG, "forceBeamColumn (3D): the section needs a torsional stiffness");
    if (r < 0) return fail(XG_ERR_ARG, "forceBeamColumn init failed");
    d->elemap[tags[i]] = std::move(e);
  }
  d->ready = false;
  return 0;
}
int orc_nodal_loads(xg_domain* d, int n, const int* tags, const double* P) {
  for (int i = 0; i < n; i++) {
    auto it = d->nodemap.find(tags[i]);
    if (it == d->nodemap.end()) return fail(XG_ERR_ARG, "load: node not found");
    for (int j = 0; j < d->ndf; j++) it->second.P[j] += P[i * d->ndf + j];
    if (d->ready) {
      auto& nr = d->nodes[d->node_index[tags[i]]];
      for (int j = 0; j < d->ndf; j++) nr.P[j] += P[i * d->ndf + j];
    }
  }
  return 0;
}
int orc_uniform_excitation(xg_domain* d, int dir, int npts, double dt, const double* accel, double factor) {
  d->gm_dir = dir;
  d->gm_dt = dt;
  d->gm_factor = factor;
  d->gm_accel.assign(accel, accel + npts);
  return 0;
}
// Domain::setRayleighDampingFactors (Domain.cpp:1858-1873) on every element
// and node.  Initial-stiffness damping (betaKinit) is not supported.
int orc_rayleigh(xg_domain* d, double alphaM, double betaK, double betaKinit, double betaKcomm) {
  if (betaKinit != 0.0) return fail(XG_ERR_ARG, "rayleigh: betaKinit (initial-stiffness damping) is not supported");
  d->alphaM = alphaM;
  d->betaK = betaK;
  d->betaKinit = 0.0;
  d->betaKcomm = betaKcomm;
  if (d->ready) rayleigh_kc_init(*d);
  return 0;
}
int orc_system(xg_domain* d, int kind, double rtol, int maxit) {
  d->system = kind;
  d->sys_rtol = rtol;
  d->sys_maxit = maxit;
  return 0;
}
int orc_numberer(xg_domain* d, int kind) {
  d->numberer = kind;
  d->ready = false;
  return 0;
}
int orc_test(xg_domain* d, int kind, double tol, int max_iter) {
  d->test_kind = kind;
  d->test_tol = tol;
  d->test_maxit = max_iter;
  return 0;
}
int orc_integrator_load_control(xg_domain* d, double dlambda) {
  d->integ = 1;
  d->lc_dlambda = d->lc_dlmin = d->lc_dlmax = dlambda;
  d->lc_spec = d->lc_last = 1.0;
  return 0;
}
int orc_integrator_disp_control(xg_domain* d, int node_tag, int dof, double incr) {
  d->integ = 2;
  d->dc_node = node_tag;
  d->dc_dof = dof - 1;
  d->dc_incr = d->dc_min = d->dc_max = incr;
  d->dc_spec = d->dc_last = 1.0;
  d->phat.clear();
  return 0;
}
int orc_integrator_newmark(xg_domain* d, double gamma, double beta) {
  d->integ = 3;
  d->nm_gamma = gamma;
  d->nm_beta = beta;
  return 0;
}
int orc_analyze(xg_domain* d, int nsteps, double dt) { return analyze(*d, nsteps, dt); }
int orc_setup(xg_domain* d) { return setup(*d); }
int orc_num_nodes(xg_domain* d) { return (int)(d->ready ? d->nodes.size() : d->nodemap.size()); }
int orc_num_elements(xg_domain* d) { return (int)d->elemap.size(); }
int orc_num_eqn(xg_domain* d) { return d->neq; }
long long orc_csr_nnz(xg_domain* d) { return (long long)d->colind.size(); }
int orc_node_eqns(xg_domain* d, int* eqn) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  for (size_t i = 0; i < d->nodes.size(); i++)
    for (int j = 0; j < d->ndf; j++) eqn[i * d->ndf + j] = d->nodes[i].eq[j];
  return 0;
}
int orc_csr_pattern(xg_domain* d, int* rowptr, int* colind) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  std::copy(d->rowptr.begin(), d->rowptr.end(), rowptr);
  std::copy(d->colind.begin(), d->colind.end(), colind);
  return 0;
}
int orc_set_trial_disp(xg_domain* d, const double* U) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  for (size_t i = 0; i < d->nodes.size(); i++)
    for (int j = 0; j < d->ndf; j++) d->nodes[i].U[j] = U[i * d->ndf + j];
  return 0;
}
int orc_get_trial_disp(xg_domain* d, double* U) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  for (size_t i = 0; i < d->nodes.size(); i++)
    for (int j = 0; j < d->ndf; j++) U[i * d->ndf + j] = d->nodes[i].U[j];
  return 0;
}
int orc_incr_disp(xg_domain* d, const double* dU) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  std::vector<double> v(dU, dU + d->neq);
  incr_disp(*d, v);
  return 0;
}
int orc_update(xg_domain* d) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  return domain_update(*d);
}
int orc_form_tangent(xg_domain* d, double* vals) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  int r = form_tangent(*d);
  if (vals) std::copy(d->Aval.begin(), d->Aval.end(), vals);
  return r;
}
int orc_form_unbalance(xg_domain* d, double* B) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  int r = form_unbalance(*d);
  if (B) std::copy(d->B.begin(), d->B.end(), B);
  return r;
}
int orc_solve(xg_domain* d, const double* B, double* X) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  std::vector<double> b(B, B + d->neq), x;
  int r = direct_solve(*d, b, x);
  if (r == 0) std::copy(x.begin(), x.end(), X);
  return r;
}
int orc_spmv(xg_domain* d, const double* x, double* y) {
  for (int i = 0; i < d->neq; i++) {
    double s = 0;
    for (int k = d->rowptr[i]; k < d->rowptr[i + 1]; k++) s += d->Aval[k] * x[d->colind[k]];
    y[i] = s;
  }
  return 0;
}
int orc_commit(xg_domain* d) {
  commit_domain(*d);
  return 0;
}
int orc_revert_to_last_commit(xg_domain* d) {
  revert_domain(*d);
  return 0;
}
// Domain::revertToStart (Domain.cpp:1941-1958): nodes and elements to their
// initial state, then one update at zero displacement with dT = 0.
int orc_revert_to_start(xg_domain* d) {
  for (auto& n : d->nodes)
    for (int j = 0; j < d->ndf; j++) n.U[j] = n.Uc[j] = n.V[j] = n.Vc[j] = n.A[j] = n.Ac[j] = 0.0;
  for (auto* e : d->eles) e->revert_to_start();
  d->lambda = 0;
  d->time = 0;
  if (d->ready) {
    const double dt_keep = d->dt;
    d->dt = 0.0;
    domain_update(*d);
    d->dt = dt_keep;
  }
  return 0;
}
int orc_set_load_factor(xg_domain* d, double lambda) {
  d->lambda = lambda;
  return 0;
}
double orc_get_load_factor(xg_domain* d) { return d->lambda; }
int orc_get_point_data(xg_domain* d, int what, double* out, int* width) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  size_t off = 0;
  int w = 0;
  for (auto* e : d->eles) {
    w = e->point_data(what, out + off);
    if (w < 0) return fail(XG_ERR_ARG, "unsupported output");
    off += (size_t)w * e->npoints();
  }
  if (width) *width = w;
  return 0;
}
int orc_element_tangent(xg_domain* d, int ele_tag, double* K, int* ndof) {
  auto it = d->elemap.find(ele_tag);
  if (it == d->elemap.end()) return fail(XG_ERR_ARG, "element not found");
  it->second->tangent(*d, K);
  *ndof = it->second->ndof();
  return 0;
}
// Threads of the element loop in an OpenMP build (the CPU baseline); returns
// the count in effect, 1 in a serial build.
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// Element::getDamp (Element.cpp:181-200) of one element: alphaM M, then
// + betaK K, + betaKcomm Kc -- the matrix the transient tangent and the
// damping force use.  Oracle-only (test infrastructure).
int orc_element_damp(xg_domain* d, int ele_tag, double* D, int* ndof) {
  auto it = d->elemap.find(ele_tag);
  if (it == d->elemap.end()) return fail(XG_ERR_ARG, "element not found");
  OElement* e = it->second.get();
  const int nd = e->ndof();
  std::vector<double> M((size_t)nd * nd, 0.0), K((size_t)nd * nd, 0.0);
  e->mass(*d, M.data());
  if (d->betaK != 0.0) e->tangent(*d, K.data());
  for (int i = 0; i < nd * nd; i++) {
    double c = 0.0;
    if (d->alphaM != 0.0) c = d->alphaM * M[i];
    if (d->betaK != 0.0) c = c + d->betaK * K[i];
    if (d->betaKcomm != 0.0) c = c + d->betaKcomm * e->Kc[i];
    D[i] = c;
  }
  *ndof = nd;
  return 0;
}
// LinearSOE::addA into the assembled CSR (SparseGenRowLinSOE.cpp:237-291)
int orc_add_a(xg_domain* d, int nblk, int bs, const double* m, const int* id) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  for (int b = 0; b < nblk; b++)
    for (int i = 0; i < bs; i++) {
      const int r = id[(size_t)b * bs + i];
      if (r < 0) continue;
      for (int j = 0; j < bs; j++) {
        const int c = id[(size_t)b * bs + j];
        if (c < 0) continue;
        const long p = csr_pos(*d, r, c);
        if (p >= d->rowptr[r + 1] || d->colind[p] != c) return fail(XG_ERR_ARG, "addA: entry outside the CSR pattern");
        d->Aval[p] += m[((size_t)b * bs + i) * bs + j];
      }
    }
  return 0;
}
int orc_element_resisting_force(xg_domain* d, int ele_tag, double* R, int* ndof) {
  auto it = d->elemap.find(ele_tag);
  if (it == d->elemap.end()) return fail(XG_ERR_ARG, "element not found");
  it->second->resist(*d, R);
  *ndof = it->second->ndof();
  return 0;
}
int orc_get_counters(xg_domain* d, xg_counters* c) {
  *c = d->counters;
  c->band_bw = -1;  // direct solve, no preconditioner
  c->band_fallbacks = 0;
  c->band_reuses = 0;
  return 0;
}
// Wall time of `reps` serial updates (the reference execution model,
// Domain.cpp:1988-2006); *ms = average per update.
int orc_time_op(xg_domain* d, int op, int reps, double* ms) {
  if (!d->ready) return fail(XG_ERR_STATE, "setup first");
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; r++) {
    if (op == XG_OP_UPDATE) domain_update(*d);
    else if (op == XG_OP_TANGENT) form_tangent(*d);
    else if (op == XG_OP_UNBALANCE) form_unbalance(*d);
    else return fail(XG_ERR_ARG, "op");
  }
  auto t1 = std::chrono::steady_clock::now();
  *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
  return 0;
}

}  // extern "C"

// =============================================================================
// Point-level drivers with the same signatures as the reference harness
// (ref_harness.cpp), so tests compare the restatement with the reference on
// identical strain paths.
// =============================================================================
extern "C" {

int orc_j2_3d_path(const double* prm, int n, const double* eps, const int* action, double* stress,
                   double* tangent) {
  J2Params p{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6]};
  J2Point s;
  s.zero();
  j2_plastic_integrator(p, 0.0, s);
  int bad = 0;
  for (int i = 0; i < n; i++) {
    if (j2_3d_set_trial_strain(p, 0.0, s, eps + 6 * i) != 0) bad++;
    j2_3d_get_stress(s, stress + 6 * i);
    j2_3d_get_tangent(s, tangent + 36 * i);
    if (action[i] == 1) j2_commit(s);
  }
  return bad;
}

int orc_j2_ps_path(const double* prm, int n, const double* eps, const int* action, double* stress,
                   double* tangent) {
  J2Params p{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6]};
  J2Point s;
  s.zero();
  j2_plastic_integrator(p, 0.0, s);
  for (int i = 0; i < n; i++) {
    j2_ps_set_trial_strain(p, 0.0, s, eps + 3 * i);
    j2_ps_get_stress(s, stress + 3 * i);
    j2_ps_get_tangent(s, tangent + 9 * i);
    if (action[i] == 1) j2_commit(s);
  }
  return 0;
}

static FiberDef make_fiber(int kind, const double* p) {
  FiberDef f;
  f.kind = kind;
  if (kind == 1) f.steel = Steel02Params{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10]};
  else if (kind == 2) f.conc = concrete02_make(p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
  else f.E = p[0];
  return f;
}

int orc_uniaxial_path(int kind, const double* prm, int n, const double* eps, const int* action,
                      double* stress, double* tangent) {
  FiberDef f = make_fiber(kind, prm);
  FiberState s;
  if (kind == 1) steel02_revert_to_start(f.steel, s.st);
  if (kind == 2) concrete02_revert_to_start(f.conc, s.cs);
  for (int i = 0; i < n; i++) {
    double sg, tg;
    fiber_set_trial(f, s, eps[i], sg, tg);
    stress[i] = sg;
    tangent[i] = tg;
    if (action[i] == 1) {
      if (kind == 1) steel02_commit(s.st);
      else if (kind == 2) concrete02_commit(s.cs);
      else s.epsP = s.eps;
    }
    if (action[i] == 2) {
      double a, b;
      FiberSection::revert_fiber(f, s, a, b);
    }
  }
  return 0;
}

int orc_fiber_section_path(int is3d, int nfib, const int* kinds, const double* prm, const double* y,
                           const double* z, const double* A, double GJ, int compute_centroid, int n,
                           const double* dfm, const int* action, double* sout, double* ksout) {
  FiberSectionDef def;
  def.is3d = is3d != 0;
  def.GJ = is3d ? GJ : -1.0;
  def.compute_centroid = compute_centroid != 0;
  for (int f = 0; f < nfib; f++) {
    FiberDef fd = make_fiber(kinds[f], prm + 12 * f);
    fd.y = y[f];
    fd.z = z ? z[f] : 0.0;
    fd.A = A[f];
    def.fibers.push_back(fd);
  }
  FiberSection sec;
  sec.init(&def);
  int order = is3d ? 4 : 2;
  for (int i = 0; i < n; i++) {
    sec.set_trial(dfm + order * i);
    for (int k = 0; k < order; k++) sout[order * i + k] = sec.s[k];
    for (int a = 0; a < order; a++)
      for (int b = 0; b < order; b++) ksout[order * order * i + order * a + b] = sec.ks(a, b);
    if (action[i] == 1) sec.commit();
    if (action[i] == 2) sec.revert();
  }
  return 0;
}

int orc_rcm(int nv, const int* ptr, const int* adj, int* order) {
  std::vector<std::vector<int>> a(nv);
  for (int v = 0; v < nv; v++) a[v].assign(adj + ptr[v], adj + ptr[v + 1]);
  for (auto& r : a) std::sort(r.begin(), r.end());
  std::vector<int> res = rcm_number(a);
  for (int i = 0; i < nv; i++) order[i] = res[i];
  return 0;
}


// Newmark transient with direct differentiation along one parameter
// direction: dsteel = (dFy, dE0, db) per Steel02 material and dconc = dfc per
// Concrete02 material, both in ascending material-tag order.  loss = sum
// over the steps of sum_i U(node_i, dof)^2; dloss = its directional
// derivative.
int orc_ddm_transient(xg_domain* d, int nsteps, double dt, const double* dsteel, const double* dconc, int nloss,
                      const int* loss_tags, int loss_dof, double* loss, double* dloss) {
  DdmDir dir;
  int k = 0;
  for (auto& kv : d->steel) {
    dir.steel[kv.first] = SteelTheta<double>{dsteel[3 * k], dsteel[3 * k + 1], dsteel[3 * k + 2]};
    k++;
  }
  k = 0;
  for (auto& kv : d->conc) dir.conc[kv.first] = dconc[k++];
  if (!d->ready) {
    int r = setup(*d);
    if (r < 0) return r;
  }
  std::vector<int> nodes;
  for (int i = 0; i < nloss; i++) {
    auto it = d->node_index.find(loss_tags[i]);
    if (it == d->node_index.end()) return fail(XG_ERR_ARG, "ddm: loss node not found");
    nodes.push_back(it->second);
  }
  if (loss_dof < 1 || loss_dof > d->ndf) return fail(XG_ERR_ARG, "ddm: loss dof out of range");
  return ddm_transient(*d, nsteps, dt, dir, nodes, loss_dof - 1, *loss, *dloss);
}

// One material trial update along a direction: committed history (pool
// order: Steel02 10 values, Concrete02 5) and strain -> trial history, with
// the forward-mode derivative for the input direction (dh, dtheta =
// (dFy, dE0, db) or dfc, dstrain).  prm as orc_uniaxial_path.  For
// Concrete02 the previous trial sig / e are taken equal to the committed
// ones (the state right after a commit).
int orc_material_jvp(int kind, const double* prm, const double* h, int kon, double strain, const double* dh,
                     const double* dtheta, double dstrain, double* hout, int* kon_out, double* dhout) {
  FiberDef F = make_fiber(kind, prm);
  if (kind == 1) {
    SteelTheta<Dual> th{Dual(F.steel.Fy, dtheta[0]), Dual(F.steel.E0, dtheta[1]), Dual(F.steel.b, dtheta[2])};
    SteelH<Dual> c{{h[0], dh[0]}, {h[1], dh[1]}, {h[2], dh[2]}, {h[3], dh[3]}, {h[4], dh[4]},
                   {h[5], dh[5]}, {h[6], dh[6]}, {h[7], dh[7]}, {h[8], dh[8]}, {h[9], dh[9]}, kon};
    SteelH<Dual> t;
    steel02_trial_t(F.steel, th, c, Dual(strain, dstrain), t);
    const Dual o[10] = {t.epsmin, t.epsmax, t.epspl, t.epss0, t.sigs0, t.epsr, t.sigr, t.eps, t.sig, t.e};
    for (int i = 0; i < 10; i++) {
      hout[i] = o[i].v;
      dhout[i] = o[i].d;
    }
    *kon_out = t.kon;
    return 0;
  }
  if (kind == 2) {
    ConcH<Dual> c{{h[0], dh[0]}, {h[1], dh[1]}, {h[2], dh[2]}, {h[3], dh[3]}, {h[4], dh[4]}};
    ConcH<Dual> t;
    concrete02_trial_t(F.conc, Dual(F.conc.fc, dtheta[0]), c, c.sig, c.e, Dual(strain, dstrain), t);
    const Dual o[5] = {t.ecmin, t.dept, t.eps, t.sig, t.e};
    for (int i = 0; i < 5; i++) {
      hout[i] = o[i].v;
      dhout[i] = o[i].d;
    }
    *kon_out = 0;
    return 0;
  }
  return fail(XG_ERR_ARG, "material_jvp: kind must be 1 (Steel02) or 2 (Concrete02)");
}


// ---- committed state in the reference's sendSelf layouts (see xara_gpu.h) ----
static void j2_points(Dom& d, std::vector<J2Point*>& pts) {
  for (auto* e : d.eles) {
    if (auto* b = dynamic_cast<OBrick*>(e))
      for (auto& p : b->pts) pts.push_back(&p);
    else if (auto* q = dynamic_cast<OQuad*>(e))
      for (auto& p : q->pts) pts.push_back(&p);
  }
}
static void fibers(Dom& d, int kind, std::vector<FiberState*>& out) {
  for (auto* e : d.eles)
    if (auto* be = dynamic_cast<OBeam*>(e))
      for (auto& sec : be->fe.secs)
        for (size_t f = 0; f < sec.fs.size(); f++)
          if (sec.def->fibers[f].kind == kind) out.push_back(&sec.fs[f]);
}

long long orc_state_size(xg_domain* d, int kind) {
  if (!d->ready && setup(*d) < 0) return XG_ERR_STATE;
  if (kind == XG_STATE_NODES) return (long long)d->nodes.size() * d->ndf * 3;
  if (kind == XG_STATE_DOMAIN) return 2;
  if (kind == XG_STATE_J2) {
    std::vector<J2Point*> p;
    j2_points(*d, p);
    return (long long)p.size() * 10;
  }
  if (kind == XG_STATE_STEEL02 || kind == XG_STATE_CONCRETE02) {
    std::vector<FiberState*> f;
    fibers(*d, kind == XG_STATE_STEEL02 ? 1 : 2, f);
    return (long long)f.size() * (kind == XG_STATE_STEEL02 ? 11 : 5);
  }
  return fail(XG_ERR_ARG, "state: unknown kind");
}

int orc_get_state(xg_domain* d, int kind, double* out) {
  const long long n = orc_state_size(d, kind);
  if (n < 0) return (int)n;
  if (kind == XG_STATE_DOMAIN) {
    out[0] = d->time;
    out[1] = d->lambda;
  } else if (kind == XG_STATE_NODES) {
    const int ndf = d->ndf;
    for (size_t k = 0; k < d->nodes.size(); k++)
      for (int j = 0; j < ndf; j++) {
        out[k * 3 * ndf + j] = d->nodes[k].Uc[j];
        out[k * 3 * ndf + ndf + j] = d->nodes[k].Vc[j];
        out[k * 3 * ndf + 2 * ndf + j] = d->nodes[k].Ac[j];
      }
  } else if (kind == XG_STATE_J2) {
    std::vector<J2Point*> p;
    j2_points(*d, p);
    for (size_t k = 0; k < p.size(); k++) {
      out[k * 10] = p[k]->xi_n;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[k * 10 + 1 + 3 * i + j] = p[k]->epsp_n[i][j];
    }
  } else if (kind == XG_STATE_STEEL02) {
    std::vector<FiberState*> f;
    fibers(*d, 1, f);
    for (size_t k = 0; k < f.size(); k++) {
      const Steel02State& s = f[k]->st;
      const double v[11] = {s.epsminP, s.epsmaxP, s.epsplP, s.epss0P, s.sigs0P, s.epssrP,
                            s.sigsrP,  (double)s.konP, s.epsP, s.sigP, s.eP};
      for (int c = 0; c < 11; c++) out[k * 11 + c] = v[c];
    }
  } else {
    std::vector<FiberState*> f;
    fibers(*d, 2, f);
    for (size_t k = 0; k < f.size(); k++) {
      const Concrete02State& s = f[k]->cs;
      const double v[5] = {s.ecminP, s.deptP, s.epsP, s.sigP, s.eP};
      for (int c = 0; c < 5; c++) out[k * 5 + c] = v[c];
    }
  }
  return 0;
}

int orc_set_state(xg_domain* d, int kind, const double* in) {
  const long long n = orc_state_size(d, kind);
  if (n < 0) return (int)n;
  if (kind == XG_STATE_DOMAIN) {
    d->time = in[0];
    d->lambda = in[1];
  } else if (kind == XG_STATE_NODES) {
    const int ndf = d->ndf;
    for (size_t k = 0; k < d->nodes.size(); k++)
      for (int j = 0; j < ndf; j++) {
        NodeRec& r = d->nodes[k];
        r.Uc[j] = r.U[j] = in[k * 3 * ndf + j];
        r.Vc[j] = r.V[j] = in[k * 3 * ndf + ndf + j];
        r.Ac[j] = r.A[j] = in[k * 3 * ndf + 2 * ndf + j];
      }
  } else if (kind == XG_STATE_J2) {
    std::vector<J2Point*> p;
    j2_points(*d, p);
    for (size_t k = 0; k < p.size(); k++) {
      p[k]->xi_n = p[k]->xi_np1 = in[k * 10];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p[k]->epsp_n[i][j] = p[k]->epsp_np1[i][j] = in[k * 10 + 1 + 3 * i + j];
    }
  } else if (kind == XG_STATE_STEEL02) {
    std::vector<FiberState*> f;
    fibers(*d, 1, f);
    for (size_t k = 0; k < f.size(); k++) {
      Steel02State& s = f[k]->st;
      const double* v = in + k * 11;
      s.epsminP = v[0];
      s.epsmaxP = v[1];
      s.epsplP = v[2];
      s.epss0P = v[3];
      s.sigs0P = v[4];
      s.epssrP = v[5];
      s.sigsrP = v[6];
      s.konP = (int)v[7];
      s.epsP = v[8];
      s.sigP = v[9];
      s.eP = v[10];
      steel02_revert(s);
    }
  } else {
    std::vector<FiberState*> f;
    fibers(*d, 2, f);
    for (size_t k = 0; k < f.size(); k++) {
      Concrete02State& s = f[k]->cs;
      const double* v = in + k * 5;
      s.ecminP = v[0];
      s.deptP = v[1];
      s.epsP = v[2];
      s.sigP = v[3];
      s.eP = v[4];
      concrete02_revert(s);
    }
  }
  return 0;
}

}  // extern "C"
