// Host object of the batched frame ensemble (see xg_ensemble.cpp).
#pragma once
#include <memory>
#include <vector>

#include "xg_domain.h"
#include "xg_frame.h"
#include "xg_ensemble.h"

namespace xg {

struct ModelDef {
  std::map<int, NodeDef> nodes;
  std::map<int, Steel02Def> steel;
  std::map<int, Concrete02Def> conc;
  std::map<int, double> elastic;
  std::map<int, FiberSecDef> secs;
  std::map<int, std::vector<double>> transfs;
  std::map<int, EleDef> eles;
  int gm_dir = -1;
  double gm_dt = 0, gm_factor = 1;
  std::vector<double> gm;
  double alphaM = 0, gamma = 0.5, beta = 0.25;
  int test_kind = XG_TEST_NORM_DISP_INCR, test_maxit = 25;
  double test_tol = 1e-8;
};

struct TuneResult {
  int block, grid, order;
  double ms;
};

struct Ensemble {
  int device = 0;
  bool ready = false, use_graph = true;
  std::vector<std::unique_ptr<ModelDef>> models;
  Domain merged;  // merged definitions + stream for the frame group
  std::unique_ptr<Group> frames;
  EnsDev E;
  int nnode = 0, neq = 0, nele = 0;
  int block = 256, grid = 0, order_mode = 0;
  long long steps_done = 0;
  std::vector<int> order_h, size_h, model_ele, node_off_h;
  DVec<int> d_node_off, d_ele_off, d_row_off, d_eq_off, d_bw, d_bent_off, d_gm_off, d_gm_n, d_eqmap, d_eq2dof, d_order;
  DVec<long long> d_band_off, d_bent_pos;
  DVec<int> d_bent_ptr, d_bent_idx, d_rhs_ptr, d_rhs_idx;
  DVec<double> d_gm_dt, d_gm_fac, d_alphaM, d_M, d_R, d_gm;
  DVec<double> U, V, A, Uc, Vc, Ac, B, X, band, Ke, Re;
  DVec<int> status, iter;
  DVec<double> time, norm;
  DVec<long long> newton_count;
  // one captured step per pool parity: the committed and trial fiber pools
  // exchange roles at every commit (flip), so the two graphs differ only in
  // which pool pointers their kernels receive
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  int parity = 0;
  double graph_dt = -1;

  ~Ensemble();
  void add(const Domain& d);
  void setup();
  int analyze(int nsteps, double dt);
  std::vector<TuneResult> tune();
  double time_config(int blk, int grd, int reps);
  void set_order(int mode);
  void enqueue_step(hipStream_t s);
  void drop_graph();
  void flip();
};

}  // namespace xg
