// Batched ensemble of 2D frame models on gfx950.
//
// Three persistent kernels, each one launch over the whole ensemble; every
// workgroup walks the models assigned to it (m = order[k], k = blockIdx.x,
// blockIdx.x + gridDim.x, ...), so a model never spans workgroups and needs
// no grid-wide synchronisation:
//   k_ens_predict  -- Newmark::newStep (Newmark.cpp:105-170): predictor,
//                     time advance, ground acceleration (PathSeries), domain
//                     update, transient unbalance; starts the Newton test.
//   k_ens_iterate  -- one NewtonRaphson iteration (NewtonRaphson.cpp:160-215)
//                     for every model still iterating: element tangents,
//                     banded assembly of c1 K + c2 C + c3 M in element order,
//                     banded Cholesky and solve (one wavefront for bands up to
//                     24), Newmark::update, fiber/section
//                     update, unbalance, CTestNormDispIncr / NormUnbalance.
//   k_ens_commit   -- Domain::commit of converged models (nodal state; the
//                     fiber pools commit by exchanging the committed and
//                     trial pool pointers); revert of models whose Newton
//                     failed (they stop, as analyze returns < 0).
#include "xg_ensemble.h"
#include "xg_frame_impl.h"

namespace xg {

template <int NT>
__device__ __forceinline__ double block_sum_t(double v, double* sh) {
  // wavefront butterfly (every lane ends with the same bits: each step adds
  // the same two operands in either order), then the NT / 64 wave sums in
  // wave order: one barrier pair instead of a log2(NT)-level LDS tree
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if constexpr (NT == 64) {
    return v;
  } else {
    const int t = threadIdx.x;
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) r += sh[w];
    __syncthreads();
    return r;
  }
}

__device__ __forceinline__ double path_accel(const EnsDev& E, int m, double t) {
  const int n = E.gm_n[m];
  const double h = E.gm_dt[m];
  if (n <= 0 || !(h > 0.0) || t < 0.0) return 0.0;
  const double incr = t / h;
  const int i1 = (int)floor(incr);
  const int i2 = i1 + 1;
  if (i2 >= n) return 0.0;
  const double* g = E.gm + E.gm_off[m];
  const double v1 = g[i1], v2 = g[i2];
  return E.gm_fac[m] * (v1 + (v2 - v1) * (incr - i1));
}

// domain update + transient unbalance of model m (all threads of the block);
// First: the step's first update (the predictor's)
template <int NT, bool First>
__device__ void model_update_and_unbalance(const EnsDev& E, int m, double ag) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long s0 = (long long)E.ele_off[m] * E.f.nip, s1 = (long long)E.ele_off[m + 1] * E.f.nip;
  if (E.row_off && E.f.rw == 20) {
    // three sections of one element per wavefront item, a 20-lane row each
    __shared__ double red[NT / 64 * 5 * 64];
    for (long long w = E.row_off[m] + wave; w < E.row_off[m + 1]; w += NT / 64)
      section_rows<false, 20, First>(E.f, w, lane, E.U, red + wave * 5 * 64);
  } else {
    for (long long sec = s0 + wave; sec < s1; sec += NT / 64) section_update<false, First>(E.f, sec, lane, E.U);
  }
  __syncthreads();
  for (int e = E.ele_off[m] + tid; e < E.ele_off[m + 1]; e += NT) element_resid<false>(E.f, e, E.Re);
  __syncthreads();
  const double alphaM = E.alphaM[m];
  for (int a = E.eq_off[m] + tid; a < E.eq_off[m + 1]; a += NT) {
    double acc = 0.0;
    for (int j = E.rhs_ptr[a]; j < E.rhs_ptr[a + 1]; j++) acc += -E.Re[E.rhs_idx[j]];
    double u = 0.0;
    const double mm = E.M[a];
    if (mm != 0.0) {
      const int d = E.eq2dof[a];
      u += -(mm * E.R[a]) * ag;
      u += -mm * E.A[d];
      if (alphaM != 0.0) u += -alphaM * mm * E.V[d];
    }
    E.B[a] = acc + u;
  }
  __syncthreads();
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Banded Cholesky L L^T = A and solve, by one 64-lane wavefront (lanes of
// the calling wave only, no barriers).  A is symmetric positive definite
// with half bandwidth bw <= BW, stored by rows: A(r, c) = Lb[r*ld + c - r + bw]
// for r - bw <= c <= r.  Right-looking: at pivot j, lane i holds row j + i
// of the active window in registers (w[c] = A(j+i, j+c), c <= i) and the
// forward-substituted right-hand side y; the pivot column is broadcast with
// readlane, the trailing window updated, and the window shifted one row
// down the wave.  Column j of L (diagonal first) overwrites band row j,
// which has left the window by then.  x holds z = L^{-1} b, then the
// solution.  Returns nonzero if a pivot is not positive.
template <int BW>
__device__ int wave_band_solve(double* __restrict__ Lb, int ld, int bw, int n, const double* __restrict__ b,
                               double* __restrict__ x) {
  const int lane = threadIdx.x & 63;
  const bool act = lane <= bw;
  double w[BW + 1];
#pragma unroll
  for (int c = 0; c <= BW; c++) w[c] = 0.0;
  double y = 0.0;
  if (act && lane < n) {
    const double* row = Lb + (long long)lane * ld + (bw - lane);
#pragma unroll
    for (int c = 0; c <= BW; c++)
      if (c <= lane) w[c] = row[c];
    y = b[lane];
  }
  int bad = 0;
  for (int j = 0; j < n; j++) {
    // row j + 1 + bw enters the window after this pivot (into lane bw):
    // lane c loads its entry c now, so the loads overlap the pivot work
    const int r = j + 1 + bw;
    const double nrow = (act && r < n) ? Lb[(long long)r * ld + lane] : 0.0;
    const double ny = r < n ? b[r] : 0.0;
    const double a00 = readlane_d(w[0], 0);
    if (!(a00 > 0.0)) bad = 1;
    const double d = sqrt(a00 > 0.0 ? a00 : 1.0);
    const double l = lane == 0 ? d : (act ? w[0] / d : 0.0);
    const double xj = readlane_d(y, 0) / d;
    if (act && j + lane < n) Lb[(long long)j * ld + lane] = l;
    if (lane == 0) x[j] = xj;
    y -= lane == 0 ? 0.0 : l * xj;
    // trailing update of the window: w[c] -= l_i l_c (entries c > i unused)
#pragma unroll
    for (int c = 1; c <= BW; c++) w[c] -= l * readlane_d(l, c);
    // shift: row j+1+i moves to lane i, columns by one
#pragma unroll
    for (int c = 0; c < BW; c++) w[c] = __shfl_down(w[c + 1], 1, 64);
    w[BW] = 0.0;
    y = __shfl_down(y, 1, 64);
#pragma unroll
    for (int c = 0; c <= BW; c++) {
      const double v = readlane_d(nrow, c);
      if (lane == bw) w[c] = v;
    }
    if (lane == bw) y = ny;
  }
  // backward substitution: x_i = (z_i - sum_k L(i+k, i) x_{i+k}) / L(i, i),
  // lane k holding x_{i+k}
  double xw = 0.0;
  double lk = (act && n - 1 + lane < n) ? Lb[(long long)(n - 1) * ld + lane] : 0.0;
  for (int i = n - 1; i >= 0; i--) {
    const double lnext = (i > 0 && act && i - 1 + lane < n) ? Lb[(long long)(i - 1) * ld + lane] : 0.0;
    double part = lane == 0 ? 0.0 : lk * xw;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    const double xi = (x[i] - part) / readlane_d(lk, 0);
    xw = __shfl_up(xw, 1, 64);
    if (lane == 1) xw = xi;
    if (lane == 0) x[i] = xi;
    lk = lnext;
  }
  return bad;
}

// Both step kernels are compiled for four waves per SIMD (128 VGPRs, a few
// spilled values): 3% more model-steps/s than the 147 / 163 VGPRs (three
// waves) the compiler picks by itself.
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_ens_predict(EnsDev E) {
  for (int k = blockIdx.x; k < E.nmodel; k += gridDim.x) {
    const int m = E.order[k];
    if (E.status[m] <= -2) continue;  // stopped after a failed step
    const int tid = threadIdx.x;
    const double g = E.gamma, b = E.beta, h = E.dt;
    const double a1 = (1.0 - g / b), a2 = h * (1.0 - 0.5 * g / b), a3 = -1.0 / (b * h), a4 = 1.0 - 0.5 / b;
    const int d0 = E.node_off[m] * E.ndf, d1 = E.node_off[m + 1] * E.ndf;
    for (int d = d0 + tid; d < d1; d += NT) {
      if (E.eqmap[d] < 0) continue;
      E.U[d] = E.Uc[d];
      E.V[d] = E.Vc[d] * a1 + E.Ac[d] * a2;
      E.A[d] = E.Ac[d] * a4 + E.Vc[d] * a3;
    }
    // E.time holds the committed time; the trial time is E.time + dt
    // (Newmark.cpp newStep: time += deltaT), formed the same way everywhere
    const double ag = path_accel(E, m, E.time[m] + h);
    __syncthreads();
    model_update_and_unbalance<NT, true>(E, m, ag);
    if (tid == 0) {
      E.iter[m] = 0;
      E.status[m] = 0;
    }
    __syncthreads();
  }
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_ens_iterate(EnsDev E) {
  __shared__ double sh[NT];
  __shared__ int s_bad;
  for (int k = blockIdx.x; k < E.nmodel; k += gridDim.x) {
    const int m = E.order[k];
    if (E.status[m] != 0) continue;
    const int tid = threadIdx.x;
    const int eq0 = E.eq_off[m], n = E.eq_off[m + 1] - eq0;
    const int bw = E.bw[m], ld = bw + 1;
    double* L = E.band + E.band_off[m];
    const double g = E.gamma, b = E.beta, h = E.dt;
    const double c2 = g / (b * h), c3 = 1.0 / (b * h * h);
    // ---- tangent: element K, then c1 K + c2 C + c3 M into the band ----
    for (int e = E.ele_off[m] + tid; e < E.ele_off[m + 1]; e += NT) element_tangent<false>(E.f, e, E.Ke);
    for (long long i = tid; i < (long long)n * ld; i += NT) L[i] = 0.0;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int q = E.bent_off[m] + tid; q < E.bent_off[m + 1]; q += NT) {
      double acc = 0.0;
      for (int j = E.bent_ptr[q]; j < E.bent_ptr[q + 1]; j++) acc += E.Ke[E.bent_idx[j]];
      E.band[E.bent_pos[q]] = acc;
    }
    __syncthreads();
    const double alphaM = E.alphaM[m];
    for (int a = tid; a < n; a += NT) {
      const double mm = E.M[eq0 + a];
      if (mm != 0.0) L[(long long)a * ld + bw] += (c3 + c2 * alphaM) * mm;
    }
    __syncthreads();
    double* x = E.X + eq0;
    const double* rhs = E.B + eq0;
    if (bw <= 24) {
      // ---- one wavefront: banded Cholesky + forward substitution, then
      //      backward substitution (the other waves wait) ----
      if (tid < 64) {
        int bad;
        if (bw <= 8) bad = wave_band_solve<8>(L, ld, bw, n, rhs, x);
        else if (bw <= 16) bad = wave_band_solve<16>(L, ld, bw, n, rhs, x);
        else bad = wave_band_solve<24>(L, ld, bw, n, rhs, x);
        if (tid == 0 && bad) s_bad = 1;
      }
      __syncthreads();
    } else {
      // ---- wide bands: left-looking banded Cholesky by columns ----
      for (int j = 0; j < n; j++) {
        if (tid == 0) {
          double s = L[(long long)j * ld + bw];
          const int k0 = j - bw > 0 ? j - bw : 0;
          for (int kk = k0; kk < j; kk++) {
            const double v = L[(long long)j * ld + (kk - j + bw)];
            s -= v * v;
          }
          if (!(s > 0.0)) s_bad = 1;
          L[(long long)j * ld + bw] = sqrt(s > 0.0 ? s : 1.0);
        }
        __syncthreads();
        const double djj = L[(long long)j * ld + bw];
        for (int r = tid; r < bw; r += NT) {
          const int i = j + 1 + r;
          if (i >= n) break;
          double s = L[(long long)i * ld + (j - i + bw)];
          const int k0 = i - bw > 0 ? i - bw : 0;
          for (int kk = k0; kk < j; kk++)
            s -= L[(long long)i * ld + (kk - i + bw)] * L[(long long)j * ld + (kk - j + bw)];
          L[(long long)i * ld + (j - i + bw)] = s / djj;
        }
        __syncthreads();
      }
      if (tid == 0) {
        for (int i = 0; i < n; i++) {
          double s = rhs[i];
          const int k0 = i - bw > 0 ? i - bw : 0;
          for (int kk = k0; kk < i; kk++) s -= L[(long long)i * ld + (kk - i + bw)] * x[kk];
          x[i] = s / L[(long long)i * ld + bw];
        }
        for (int i = n - 1; i >= 0; i--) {
          double s = x[i];
          const int k1 = i + bw < n - 1 ? i + bw : n - 1;
          for (int kk = i + 1; kk <= k1; kk++) s -= L[(long long)kk * ld + (i - kk + bw)] * x[kk];
          x[i] = s / L[(long long)i * ld + bw];
        }
      }
      __syncthreads();
    }
    if (s_bad) {
      if (tid == 0) E.status[m] = -1;  // BadLinearSolve
      __syncthreads();
      continue;
    }
    // ---- Newmark::update ----
    const int d0 = E.node_off[m] * E.ndf, d1 = E.node_off[m + 1] * E.ndf;
    for (int d = d0 + tid; d < d1; d += NT) {
      const int a = E.eqmap[d];
      if (a < 0) continue;
      const double du = E.X[a];
      E.U[d] += du;
      E.V[d] += c2 * du;
      E.A[d] += c3 * du;
    }
    __syncthreads();
    const double ag = path_accel(E, m, E.time[m] + h);
    model_update_and_unbalance<NT, false>(E, m, ag);
    // ---- convergence test ----
    double v = 0.0;
    const double* w = E.test_kind == 1 ? E.B : E.X;
    for (int a = eq0 + tid; a < eq0 + n; a += NT) v += w[a] * w[a];
    const double norm = sqrt(block_sum_t<NT>(v, sh));
    if (tid == 0) {
      const int it = E.iter[m] + 1;
      E.iter[m] = it;
      E.norm[m] = norm;
      E.newton_count[m] += 1;
      if (norm <= E.tol) E.status[m] = 1;
      else if (it >= E.maxit) E.status[m] = -1;
    }
    __syncthreads();
  }
}

// commit converged models; revert (and stop) models whose step failed
template <int NT>
__global__ __launch_bounds__(NT) void k_ens_commit(EnsDev E) {
  for (int k = blockIdx.x; k < E.nmodel; k += gridDim.x) {
    const int m = E.order[k];
    const int st = E.status[m];
    if (st <= -2) continue;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int d0 = E.node_off[m] * E.ndf, d1 = E.node_off[m + 1] * E.ndf;
    const long long s0 = (long long)E.ele_off[m] * E.f.nip, s1 = (long long)E.ele_off[m + 1] * E.f.nip;
    if (st == 1) {
      if (tid == 0) E.time[m] += E.dt;
      for (int d = d0 + tid; d < d1; d += NT) {
        E.Uc[d] = E.U[d];
        E.Vc[d] = E.V[d];
        E.Ac[d] = E.A[d];
      }
      // the fiber state commits by the host exchanging the committed and
      // trial pools after this kernel (Ensemble::flip)
    } else {
      // failed (or never converged within the launched iterations):
      // Domain::revertToLastCommit and stop this model
      for (int d = d0 + tid; d < d1; d += NT) {
        E.U[d] = E.Uc[d];
        E.V[d] = E.Vc[d];
        E.A[d] = E.Ac[d];
      }
      for (long long sec = s0 + wave; sec < s1; sec += NT / 64) section_revert<false>(E.f, sec, lane);
      __syncthreads();
      if (tid == 0) E.status[m] = -2;
    }
    __syncthreads();
  }
}

#define XG_ENS_DISPATCH(KERNEL, ...)                                                                    \
  switch (block) {                                                                                     \
    case 128: hipLaunchKernelGGL(KERNEL<128>, dim3(grid), dim3(128), 0, s, __VA_ARGS__); break;         \
    case 512: hipLaunchKernelGGL(KERNEL<512>, dim3(grid), dim3(512), 0, s, __VA_ARGS__); break;         \
    default: hipLaunchKernelGGL(KERNEL<256>, dim3(grid), dim3(256), 0, s, __VA_ARGS__); break;          \
  }

void launch_ens_predict(hipStream_t s, const EnsDev& E, int grid, int block) { XG_ENS_DISPATCH(k_ens_predict, E); }
void launch_ens_iterate(hipStream_t s, const EnsDev& E, int grid, int block) { XG_ENS_DISPATCH(k_ens_iterate, E); }
void launch_ens_commit(hipStream_t s, const EnsDev& E, int grid, int block) { XG_ENS_DISPATCH(k_ens_commit, E); }

}  // namespace xg
