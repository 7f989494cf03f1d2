// Discrete adjoint of a Newmark transient of a frame model (dispBeamColumn
// over Steel02 / Concrete02 / Elastic fiber sections): gradients of a
// displacement loss with respect to the Steel02 (Fy, E0, b) and Concrete02
// fc of every material, with the committed state checkpointed every k steps
// and recomputed in between.
#pragma once
#include <vector>

#include "xg_domain.h"
#include "xg_frame.h"
#include "xg_frame_dev.h"

namespace xg {

struct AdjCoef {
  double c2, c3, a1, a2, a3, a4, alphaM;
};

void launch_adj_strain_adjoint(hipStream_t s, const FrameDev& f, bool d3, const double* U, const double* hbs,
                               const double* hbc, double* shat);
void launch_adj_history(hipStream_t s, const FrameDev& f, bool d3, const double* U, const double* mu, double* hbs,
                        double* hbc, double* gs, double* gc);
void launch_adj_rhs(hipStream_t s, int n, const double* ub, const double* vb, const double* ab, const double* neg_g,
                    double c2, double c3, double* rhs);
void launch_adj_newmark(hipStream_t s, int n, const double* M, const double* mu, const AdjCoef& k, double* ub,
                        double* vb, double* ab);
void launch_loss_grad(hipStream_t s, int nl, const int* dofs, const int* eqs, const double* U, double* ub);
void launch_loss_accum(hipStream_t s, int nl, const int* dofs, const double* U, double* acc);
void launch_reduce_mat(hipStream_t s, int nmat, const int* ptr, const int* idx, const double* vals, long long npool,
                       int ncomp, double* out);

struct Adjoint {
  // loss = sum over steps of sum_i U(loss_dof_i)^2 (nodal DOF indices)
  std::vector<int> loss_dof_h;
  DVec<int> loss_dof, loss_eq;
  // last forward run
  bool have_forward = false;
  int nsteps = 0, ckpt = 0, nck = 0;
  double dt = 0, loss = 0, fwd_ms = 0, bwd_ms = 0;
  long long bwd_cg = 0;
  long long nst = 0, ncs = 0, nel = 0, ndof = 0;
  std::vector<double> ck_time;
  DVec<double> ck_st, ck_cs, ck_el, ck_U, ck_V, ck_A;  // nck checkpoints + the final state
  DVec<int> ck_kon;
  DVec<double> seg_st, seg_cs, seg_el, seg_U;          // one segment, per step
  DVec<int> seg_kon;
  DVec<double> acc;
  // reverse sweep
  DVec<double> ub, vb, ab, rhs, mu, neg_g, mu_nodal, shat, hbs, hbc, gs, gc, grad_s, grad_c;
  DVec<int> smat_ptr, smat_idx, cmat_ptr, cmat_idx;
  int nsmat = 0, ncmat = 0;

  void set_loss(Domain& d, const std::vector<int>& node_tags, int dof);
  double forward(Domain& d, int nsteps, double dt, int ckpt);
  void backward(Domain& d, double* grad_steel, double* grad_conc);

 private:
  void check(Domain& d);
  void save_pools(Domain& d, double* st, int* kon, double* cs, double* el);
  void load_pools(Domain& d, const double* st, const int* kon, const double* cs, const double* el);
  void save_ck(Domain& d, int slot);
  void load_ck(Domain& d, int slot);
  void material_maps(Domain& d);
};


}  // namespace xg
