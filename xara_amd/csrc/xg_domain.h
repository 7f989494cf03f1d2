// One finite-element model resident on one GPU (the xg_domain handle).
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/xara_gpu.h"
#include "xg_common.h"
#include "xg_kernels.h"
#include "xg_memory.h"
#include "xg_numbering.h"

namespace xg {

struct Domain;

// A group of elements of one formulation, stored in ascending tag order.
struct Group {
  int nen = 0;                 // nodes per element
  int nd = 0;                  // element DOFs
  std::vector<int> tags;       // ascending
  std::vector<int> conn;       // node indices, nele*nen
  long long ke_off = 0, re_off = 0;
  virtual ~Group() {}
  int nele() const { return (int)tags.size(); }
  virtual int npts() const = 0;  // material points (or sections) per element
  virtual long long fibers() const { return (long long)nele() * npts(); }
  virtual void setup(Domain& d) = 0;
  virtual void update(Domain& d) = 0;
  virtual void resid(Domain& d, double* Re) = 0;
  virtual void tangent(Domain& d, double* Ke) = 0;
  // formTangent of a model made of this group alone straight into the CSR
  // values, where the group has such a kernel; false if it has not
  virtual bool tangent_csr(Domain&) { return false; }
  virtual void mass_diag(Domain&, std::vector<double>&) {}  // lumped element mass per element DOF
  virtual void commit(Domain& d) = 0;
  virtual void revert(Domain& d) = 0;
  virtual void revert_to_start(Domain& d) = 0;
  // per-point outputs for element i (local index) appended to out; returns width
  virtual int point_data(Domain& d, int what, std::vector<double>& out) = 0;
  virtual size_t device_bytes() const { return 0; }
  // committed material state in the reference's sendSelf layout (XG_STATE_*
  // kinds of the C-ABI): number of doubles, read, write (committed and trial)
  virtual long long state_size(int) const { return 0; }
  virtual void get_state(Domain&, int, double*) {}
  virtual void set_state(Domain&, int, const double*) {}
  // an element update that failed since the last call (Element::update < 0)
  virtual bool update_failed(Domain&) { return false; }
};

struct NodeDef {
  double crd[3] = {0, 0, 0};
  int fix[6] = {0, 0, 0, 0, 0, 0};
  double mass[6] = {0, 0, 0, 0, 0, 0};
  double P[6] = {0, 0, 0, 0, 0, 0};
};

struct EleDef {
  int type = 0;  // 1 brick, 2 quad, 3 dispbeam, 4 forcebeam
  std::vector<int> nodes;  // node tags
  int mat = 0, sec = 0, transf = 0, nip = 0, maxiter = 10;
  double thk = 1.0, rho = 0.0, b[3] = {0, 0, 0}, tol = 1e-12;
};

struct Steel02Def {
  double Fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sigini;
};
struct Concrete02Def {
  double fc, epsc0, fcu, epscu, rat, ft, Ets;
};
struct FiberSecDef {
  std::vector<int> kind;  // 1 steel02, 2 concrete02, 3 elastic
  std::vector<int> mat;   // material tag
  std::vector<double> y, z, A;
  double GJ = -1.0;
  bool centroid = true;
};

struct Domain {
  int ndm = 0, ndf = 0, device = 0;
  hipStream_t stream = nullptr;
  // ---- definition ----
  std::map<int, NodeDef> nodedefs;
  std::map<int, J2Param> j2defs;
  std::map<int, Steel02Def> steeldefs;
  std::map<int, Concrete02Def> concdefs;
  std::map<int, double> elasticdefs;
  std::map<int, FiberSecDef> secdefs;
  std::map<int, std::vector<double>> transfdefs;
  std::map<int, EleDef> eledefs;
  // ---- after setup ----
  bool ready = false;
  int nnode = 0;
  std::vector<int> node_tags;
  std::map<int, int> node_index;
  std::vector<double> crd_h;  // nnode*ndm
  Numbering num;
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<std::pair<int, int>> ele_order;  // ascending tag -> (group, local)
  long long ke_total = 0, re_total = 0;
  // device model
  DVec<double> crd, U, Uc, V, Vc, A, Ac;
  DVec<int> eqmap;
  DVec<int> rowptr, colind;
  // node-blocked column indices of the CSR pattern (k_spmv_nb_dot), when
  // every node has all ndf DOFs free or none: nb_m = ndf, else 0
  DVec<int> nbcol;
  int nb_m = 0;
  // block-symmetric copy for the SpMV (XG_SPMV_SYM), built on first use and
  // refreshed from the CSR values after each assembly
  DVec<int> sym_ptr, sym_col, sym_blk, sym_src, sym_len;
  DVec<double> sym_val;
  int sym_nup = 0;
  bool sym_built = false, sym_ok = false, sym_stale = true;
  SymSpmv sym_view;
  // the copy's block rows in the Morton order of their nodes (sym_sweep):
  // the chunked PCG then runs in that order -- b and a warm start's guess
  // gathered into sw_b / sw_x, the block-Jacobi inverses into binv_s, the
  // solution scattered back (sym_ord[position] = block row)
  bool sym_sweep = false;
  DVec<int> sym_ord;
  DVec<double> sw_b, sw_x, binv_s;
  DVec<double> vals, Ke, Re, B, X, Peq;
  DVec<int> asm_ptr, asm_idx, rhs_ptr, rhs_idx;
  DVec<int> nfail, pcbad;
  // solver workspace
  DVec<double> r0, r1, z, p, p2, q, part, binv, tmpv;
  DVec<int> bfirst, blk_first;
  DVec<unsigned char> bsize, blk_size;
  // stdBrick-only models: per node block row, the elements around the node
  // (ascending tag), the node's local index in each, and each element
  // node's neighbour position in the row (k_brick_tangent_csr)
  DVec<int> fz_ptr, fz_ele;
  DVec<unsigned char> fz_a, fz_tgt;
  bool fz_ok = false;
  void fused_tangent_tables();
  DVec<char> scal;
  DVec<double> dscalar;
  int nblk = 0;
  // analysis configuration
  int system = XG_SYSTEM_PCG_JACOBI, numberer = XG_NUMBERER_RCM;
  double rtol = 1e-10;
  int maxit = 20000;
  int spmv_lanes = 0;  // 0 = choose from the mean row length
  int spmv_format = XG_SPMV_AUTO;
  int cg_chunk = 16;  // PCG iterations issued between convergence reads (even)
  // full chunks replayed from captured graphs (kernels only), one per set of
  // launch arguments (the solution vector differs between the solves of
  // DisplacementControl, for example); kept until teardown or cg_config
  struct CgGraph {
    std::vector<uintptr_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  static constexpr int kCgGraphs = 16;
  bool cg_graph_on = true;
  bool cg_fuse_on = true;  // alpha / beta reductions fused into their neighbours (small systems)
  bool cg_small_on = true; // whole solve in one workgroup where the system fits (cg_small_fits)
  // warm starts of the PCG: under LoadControl the first Newton iteration of a
  // step starts from the previous step's first solution, DisplacementControl's
  // reference-load solves from the previous reference-load solution (the
  // chunked PCG only; a solve still stops at |r| <= rtol |b|)
  bool cg_warm_on = true;
  DVec<double> Xw;          // the previous step's first-iteration solution
  bool xw_valid = false, dc_hat_valid = false;
  std::vector<CgGraph> cg_graphs;
  int test_kind = XG_TEST_NORM_UNBALANCE;
  double test_tol = 1e-6;
  int test_maxit = 25;
  int integ = 1;
  double lc_dlambda = 1, lc_min = 1, lc_max = 1, lc_spec = 1, lc_last = 1;
  int dc_node = 0, dc_dof = 0, dc_eq = -1;
  double dc_incr = 0, dc_min = 0, dc_max = 0, dc_spec = 1, dc_last = 1;
  bool dc_ready = false;
  DVec<double> phat, dUhat, dUbar, dU;
  double lambda = 0, time = 0, dt = 0;
  // transient (Newmark, UniformExcitation + PathSeries, Rayleigh damping)
  double nm_gamma = 0.5, nm_beta = 0.25, nm_c2 = 0, nm_c3 = 0;
  double alphaM = 0;
  // stiffness-proportional Rayleigh factors on the current tangent (betaK)
  // and on the tangent of the last commit (betaKc, stored per element in
  // Kc with the layout of Ke, Element.cpp:74-80,167-176); ray_conn: each
  // group's connectivity on the device for the damping force
  double betaK = 0, betaKc = 0;
  DVec<double> Kc;
  std::vector<std::unique_ptr<DVec<int>>> ray_conn;
  bool stiffness_damping() const { return integ == 3 && (betaK != 0.0 || betaKc != 0.0); }
  void rayleigh_kc_init();
  void rayleigh_forces();
  int gm_dir = -1;
  double gm_dt = 0, gm_factor = 1, ag = 0;
  std::vector<double> gm_accel;
  DVec<double> Meq, Req, mass_val;
  DVec<int> eq2dof;
  DVec<long long> mass_pos;
  int nmass_eq = 0;
  void setup_transient();
  double ground_accel(double t) const;
  xg_counters counters{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  ~Domain();
  // operations
  void setup();
  void update();
  void form_unbalance();
  void form_tangent();
  // the buffers one PCG solve works in: the domain's own (cg_bufs()) or a
  // caller's workspace of the same sizes
  struct CgBufs {
    double *r0, *r1, *z, *p, *p2, *q, *part, *binv;
    int* pcbad;
    void* scal;
    double *bandA, *bandL, *bandLr;  // band matrix and its Cholesky factor (band_active())
    double *bs = nullptr, *xs = nullptr, *binvs = nullptr;  // b, x and binv in the sweep order (sym_sweep)
    bool wide = false;  // the domain's wide-band factor may be used (wide_active())
  };
  // the symmetric product in its sweep order runs this solve (chunked PCG)
  bool sweep_active(const CgBufs& w) { return sym_sweep && w.bs && w.xs && w.binvs && sym_prepare(true); }
  // band-Cholesky preconditioner of small systems (taken by the one-
  // workgroup solve, cg_small): the half bandwidth in the band order
  // (the numbering, or the nodes sorted along the model's longest extent
  // when that is narrower), -1 when neither is 63 wide or less or the
  // system is larger than kBandMaxEqn
  static constexpr int kBandMaxEqn = 16384;
  int band_bw = -1;
  DVec<int> band_pinv;
  DVec<double> band_A, band_L, band_Lr;
  DVec<double> band2;  // work vectors of a second band solve (solve_pair)
  DVec<char> scal2;
  bool dc_hat_ready = false;  // the Newton iteration has solved for the reference load
  int solve_pair(const double* b1, double* x1, const double* b2, double* x2);
  bool band_active() const { return cg_small_on && band_bw >= 0 && !wb_failed; }
  void band_order();
  // band-Cholesky preconditioner of wider bands (xg_wband.hip): systems of
  // more than 1,024 equations (the one-workgroup solve takes smaller ones)
  // whose half bandwidth in the band order is at most kWbMaxBw, or at most
  // 63 with more than kBandMaxEqn equations, and whose factor fits in
  // kWbMaxBytes, under the static integrators (a transient's K + c3 M is
  // well conditioned by the mass term: block Jacobi converges in tens of
  // iterations there, faster than a factorisation per Newton iteration).
  // The PCG preconditioned by the exact factor converges in one or two
  // iterations (solve_wide); factor and sweeps are replayed from captured
  // graphs when cg_graph_on.
  static constexpr int kWbMaxEqn = 1 << 18;
  static constexpr size_t kWbMaxBytes = size_t(1) << 30;
  int wb_bw = -1;
  DVec<double> wb_L, wb_D, wb_Y, wb_v;  // factor tiles, diagonal tiles, 3 x 2 sweep vectors, 4 PCG vectors
  hipGraph_t wb_fg = nullptr, wb_sg[2] = {nullptr, nullptr};
  hipGraphExec_t wb_fe = nullptr, wb_se[2] = {nullptr, nullptr};
  bool cg_wide_on = true;
  // the current matrix's band factor (either kind) met a pivot that is not
  // positive (an indefinite tangent past a peak): its solves fall back to
  // the PCG without it until the next assembly
  bool wb_failed = false;
  long long wb_fallbacks = 0;
  // the values of the matrix whose band factor (either kind) the domain's
  // own factor storage holds: a matrix equal to them bit for bit (a model
  // still elastic) keeps its factor instead of being factored again
  // (XG_PC_REUSE=0 turns this off)
  DVec<double> pc_vals;
  DVec<int> pc_diff;
  bool pc_reuse_on = true, pc_cache = false, pc_fresh = false;
  const void* pc_cache_key = nullptr;
  const void* pc_fresh_key = nullptr;
  long long pc_reuses = 0;
  bool wide_active() const {
    return cg_small_on && cg_wide_on && !wb_failed && wb_bw >= 0 && band_bw < 0 && integ != 3;
  }
  void wb_factor(int* bad);
  void pc_buffers();
  void wb_apply(int k, const double* const* r, double* const* z, bool graphs);
  // warm: bit s set when x[s] holds a guess to start right-hand side s from
  int solve_wide(int R, const double* const* b, double* const* x, const CgBufs& w, bool graphs, unsigned warm = 0);
  // the preconditioner's buffers last formed from the current matrix
  // (nullptr: the matrix changed since), so repeated solves of one matrix
  // form it once
  const void* pc_done = nullptr;
  CgBufs cg_bufs();
  int solve(const double* b, double* x) { return solve_with(b, x, cg_bufs(), cg_graph_on); }  // device pointers
  // as solve, starting from the guess in x (warm) or from zero
  int solve_from(const double* b, double* x, bool warm) { return solve_with(b, x, cg_bufs(), cg_graph_on, warm); }
  // returns < 0 on failure; graphs: replay chunks from captured graphs
  int solve_with(const double* b, double* x, const CgBufs& w, bool graphs, bool warm = false);
  bool precond_ready(const CgBufs& w);
  void incr_disp(const double* dU_dev, double scale = 1.0);
  void commit();
  void revert();
  void revert_to_start();
  double norm2_dev(const double* v, int n);
  double dot_dev(const double* a, const double* b, int n);
  double read_eq(const double* v, int a);
  int analyze(int nsteps, double dtv);
  long long state_size(int kind);
  void get_state(int kind, double* out);
  void set_state(int kind, const double* in);
  int newton();
  int new_step();
  int integ_update(const double* x);
  int update_checked();  // update(); < 0 if an element update failed (AnalysisModel::updateDomain)
  int run_test(int& iter, std::vector<double>& norms, int& numIncr);
  void precond_setup(const CgBufs& w);
  void cg_chunk_run(int iters, int it0, int L, int bw, double* x, const CgBufs& w, bool graphs);
  void cg_graph_clear();
  int lanes();
  std::vector<int> node_block_columns();
  int spmv_nbm() const { return spmv_format == XG_SPMV_CSR ? 0 : nb_m; }
  // the symmetric view to pass to the SpMV launches (nullptr when another
  // layout is in use); converts stale values first.  solver: for the PCG,
  // which assumes a symmetric matrix and takes it under XG_SPMV_AUTO too
  const SymSpmv* sym_prepare(bool solver);
  // y = K x (device pointers) in the configured layout (solver: as the PCG
  // forms it)
  void spmv(const double* x, double* y, bool solver = false);
  // the PCG's product on vectors in its own order (the sweep order when
  // sym_sweep, else the equation order)
  void spmv_sweep(const double* x, double* y);
  // the whole solve runs in one workgroup (k_pcg_small)
  bool small_path() const { return cg_small_on && cg_small_fits(num.neq, (long long)num.colind.size()); }
};

// no-allocation solve (xg_noalloc.cpp): per-device scratch lookup at
// xg_create, and the solve with its work buffers in the scratch array
void noalloc_init(int device);
size_t noalloc_scratch_bytes();
int solve_noalloc(Domain& d, const double* B, double* X);

}  // namespace xg
