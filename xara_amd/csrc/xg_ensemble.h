// Batched ensemble of independent 2D frame models (dispBeamColumn2d over
// fiber sections) advanced in lock step by Newmark + Newton: one persistent
// launch per Newton iteration, every workgroup owning whole models.
#pragma once
#include "xg_frame_dev.h"

namespace xg {

struct EnsDev {
  int nmodel = 0;
  FrameDev f;  // all elements of all models; node indices are global
  int ndf = 3;
  // per model (size nmodel or nmodel+1)
  const int* node_off = nullptr;
  const int* ele_off = nullptr;
  const int* row_off = nullptr;   // first wavefront item of the frame rows (f.rsec), nullptr: none
  const int* eq_off = nullptr;
  const long long* band_off = nullptr;
  const int* bw = nullptr;
  const int* bent_off = nullptr;  // band-entry range of each model
  const int* gm_off = nullptr;
  const int* gm_n = nullptr;
  const double* gm_dt = nullptr;
  const double* gm_fac = nullptr;
  const double* alphaM = nullptr;
  const int* order = nullptr;     // processing order of models (autotuned)
  // per global node DOF / equation
  const int* eqmap = nullptr;     // node DOF -> global equation, -1 fixed
  const int* eq2dof = nullptr;    // global equation -> node DOF
  const double* M = nullptr;      // nodal mass per equation
  const double* R = nullptr;      // influence vector per equation
  const double* gm = nullptr;     // concatenated ground-acceleration records
  double *U = nullptr, *V = nullptr, *A = nullptr, *Uc = nullptr, *Vc = nullptr, *Ac = nullptr;
  double *B = nullptr, *X = nullptr;
  double* band = nullptr;
  double *Ke = nullptr, *Re = nullptr;
  // band assembly (lower triangle in the model's CSR pattern) and rhs maps
  const long long* bent_pos = nullptr;  // global band index of each entry
  const int* bent_ptr = nullptr;
  const int* bent_idx = nullptr;        // Ke entries, element order
  const int* rhs_ptr = nullptr;
  const int* rhs_idx = nullptr;         // Re entries per global equation
  // dynamic state per model
  int* status = nullptr;   // 0 running, 1 converged this step, -1 failed this step, -2 stopped
  int* iter = nullptr;
  double* time = nullptr;  // committed time of each model
  double* norm = nullptr;
  long long* newton_count = nullptr;
  // analysis parameters
  double gamma = 0.5, beta = 0.25, dt = 0.01, tol = 1e-8;
  int maxit = 25, test_kind = 2;
};

void launch_ens_predict(hipStream_t s, const EnsDev& E, int grid, int block);
void launch_ens_iterate(hipStream_t s, const EnsDev& E, int grid, int block);
void launch_ens_commit(hipStream_t s, const EnsDev& E, int grid, int block);

}  // namespace xg
