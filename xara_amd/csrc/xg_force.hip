// ForceBeamColumn2d / ForceBeamColumn3d state determination on gfx950.
//
// One 256-thread workgroup per element.  The element-level compatibility
// iteration of ForceBeamColumn2d::update (ForceBeamColumn2d.cpp:559-934;
// ForceBeamColumn3d.cpp:588-1040) runs with the section loop spread over the
// four wavefronts (wave w takes sections w, w + 4, ...; each section's fibers
// are lane-strided as in the displacement-based update) and the element
// algebra -- flexibility sum, its inverse, the residual deformation and the
// energy test -- on thread 0, with the iteration state in LDS.  Every
// decision (accept, next algorithm, subdivide) is taken by thread 0 and read
// by the whole workgroup after a barrier, so the control flow stays uniform.
// Sections are integrated by the reference's Gauss-Lobatto rule (locations
// and weights uploaded per group); no element loads.
#include <cfloat>

#include "../../include/xara_gpu.h"
#include "xg_frame_impl.h"

namespace xg {

constexpr int kForceMaxSec = 10;
constexpr int kForceThreads = 256;

// Gauss-Jordan inverse with partial pivoting (the first largest pivot) of a
// row-major N x N matrix; every index is a compile-time constant so the
// matrix stays in registers.  Returns false on a zero pivot.
template <int N>
__device__ __forceinline__ bool inverse_n(const double* A, double* Ai) {
  double a[N][2 * N];
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = 0; j < 2 * N; j++) a[i][j] = j < N ? A[i * N + j] : (j - N == i ? 1.0 : 0.0);
#pragma unroll
  for (int c = 0; c < N; c++) {
    int piv = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int i = c + 1; i < N; i++) {
      const double v = fabs(a[i][c]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    if (best == 0.0) return false;
#pragma unroll
    for (int i = c + 1; i < N; i++) {
      const bool sw = piv == i;
#pragma unroll
      for (int j = 0; j < 2 * N; j++) {
        const double x = a[c][j], y = a[i][j];
        a[c][j] = sw ? y : x;
        a[i][j] = sw ? x : y;
      }
    }
    const double inv = 1.0 / a[c][c];
#pragma unroll
    for (int j = 0; j < 2 * N; j++) a[c][j] *= inv;
#pragma unroll
    for (int i = 0; i < N; i++)
      if (i != c) {
        const double fct = a[i][c];
#pragma unroll
        for (int j = 0; j < 2 * N; j++) a[i][j] -= fct * a[c][j];
      }
  }
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = 0; j < N; j++) Ai[i * N + j] = a[i][N + j];
  return true;
}

template <bool D3>
struct ForceShared {
  static constexpr int NB = D3 ? 6 : 3, NS = D3 ? 4 : 2;
  double SeT[NB], kvT[NB * NB], dvT[NB], dvToDo[NB], vin[NB];
  double vs[kForceMaxSec][NS], fs[kForceMaxSec][NS * NS], Ssr[kForceMaxSec][NS];
  double fc[kForceMaxSec][NB * NB], vr[kForceMaxSec][NB];
  int action, converged, nsub;
};

// Section i's part of one iteration: the deformation increment from the
// section force residual, the fiber update at the new deformation, its
// flexibility, and its contributions b^T fs b wtL to the element flexibility
// and b^T (vs + dvs) wtL to the residual deformation (codes P, MZ, MY, T).
template <bool D3>
__device__ __forceinline__ void force_section(const FrameDev& f, ForceShared<D3>& S, int e, int i, int lane,
                                              bool initial_fs, bool add_dvs, double L) {
  constexpr int NB = ForceShared<D3>::NB, NS = ForceShared<D3>::NS;
  const long long sec = (long long)e * f.nip + i;
  const double xL = f.lxi[i], xL1 = xL - 1.0, wtL = f.lwt[i] * L;
  double Ss[NS], dSs[NS], dvs[NS], fsl[NS * NS];
  Ss[0] = S.SeT[0];
  Ss[1] = xL1 * S.SeT[1] + xL * S.SeT[2];
  if constexpr (D3) {
    Ss[2] = xL1 * S.SeT[3] + xL * S.SeT[4];
    Ss[3] = S.SeT[5];
  }
#pragma unroll
  for (int a = 0; a < NS; a++) dSs[a] = Ss[a] - S.Ssr[i][a];
  const double* fs0 = f.fs0 + (long long)f.secdef[e] * NS * NS;
#pragma unroll
  for (int k = 0; k < NS * NS; k++) fsl[k] = initial_fs ? fs0[k] : S.fs[i][k];
#pragma unroll
  for (int a = 0; a < NS; a++) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < NS; b++) acc += fsl[a * NS + b] * dSs[b];
    dvs[a] = acc;
  }
  double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < NS; a++) d[a] = add_dvs ? S.vs[i][a] + dvs[a] : S.vs[i][a];
  double sr[4], kt[16], fsn[NS * NS];
  section_fibers<D3>(f, sec, lane, d, sr, kt);
  double ks[NS * NS];
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = 0; b < NS; b++) ks[a * NS + b] = kt[a * (D3 ? 4 : 2) + b];
  if (!inverse_n<NS>(ks, fsn))
#pragma unroll
    for (int k = 0; k < NS * NS; k++) fsn[k] = 0.0;
#pragma unroll
  for (int a = 0; a < NS; a++) dSs[a] = Ss[a] - sr[a];
#pragma unroll
  for (int a = 0; a < NS; a++) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < NS; b++) acc += fsn[a * NS + b] * dSs[b];
    dvs[a] = acc;
  }
  // fb = fs b wtL (NS x NB) and the section's rows of b^T fb, in the order
  // of the reference's code switch (P, MZ, then MY, T in 3D); fb, fc and vr
  // are 6 wide so that the 3D-only columns exist in 2D too (unused there)
  double fb[NS][6];
#pragma unroll
  for (int a = 0; a < NS; a++)
#pragma unroll
    for (int b = 0; b < 6; b++) fb[a][b] = 0.0;
#pragma unroll
  for (int ii = 0; ii < NS; ii++)
#pragma unroll
    for (int jj = 0; jj < NS; jj++) {
      const double tmp = fsn[jj * NS + ii] * wtL;
      if (ii == 0) fb[jj][0] += tmp;
      else if (ii == 1) {
        fb[jj][1] += xL1 * tmp;
        fb[jj][2] += xL * tmp;
      } else if (ii == 2) {
        fb[jj][3] += xL1 * tmp;
        fb[jj][4] += xL * tmp;
      } else
        fb[jj][5] += tmp;
    }
  if (lane == 0) {
    double fc[6][6], vr[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      vr[r] = 0.0;
#pragma unroll
      for (int c = 0; c < 6; c++) fc[r][c] = 0.0;
    }
#pragma unroll
    for (int ii = 0; ii < NS; ii++)
#pragma unroll
      for (int jj = 0; jj < NB; jj++) {
        const double tmp = fb[ii][jj];
        if (ii == 0) fc[0][jj] = tmp;
        else if (ii == 1) {
          fc[1][jj] = xL1 * tmp;
          fc[2][jj] = xL * tmp;
        } else if (ii == 2) {
          fc[3][jj] = xL1 * tmp;
          fc[4][jj] = xL * tmp;
        } else
          fc[5][jj] = tmp;
      }
#pragma unroll
    for (int ii = 0; ii < NS; ii++) {
      const double dei = (dvs[ii] + d[ii]) * wtL;
      if (ii == 0) vr[0] = dei;
      else if (ii == 1) {
        vr[1] = xL1 * dei;
        vr[2] = xL * dei;
      } else if (ii == 2) {
        vr[3] = xL1 * dei;
        vr[4] = xL * dei;
      } else
        vr[5] = dei;
    }
#pragma unroll
    for (int r = 0; r < NB; r++)
#pragma unroll
      for (int c = 0; c < NB; c++) S.fc[i][r * NB + c] = fc[r][c];
#pragma unroll
    for (int k = 0; k < NB; k++) S.vr[i][k] = vr[k];
#pragma unroll
    for (int a = 0; a < NS; a++) {
      S.vs[i][a] = d[a];
      S.Ssr[i][a] = sr[a];
    }
#pragma unroll
    for (int k = 0; k < NS * NS; k++) S.fs[i][k] = fsn[k];
  }
}

template <bool D3>
__global__ __launch_bounds__(kForceThreads) void k_force_update(FrameDev f, const double* __restrict__ U,
                                                                int* __restrict__ err) {
  constexpr int NB = ForceShared<D3>::NB, NS = ForceShared<D3>::NS;
  __shared__ ForceShared<D3> S;
  const int e = blockIdx.x;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int nip = f.nip, nw = kForceThreads / 64;
  const double L = f.geo[(long long)e * 10];
  double* Se = f.Se + (long long)e * NB;
  double* kv = f.kv + (long long)e * NB * NB;
  double* vprev = f.vprev + (long long)e * NB;
  const bool init0 = f.init[e] != 0;
  if (t == 0) {
    double v[6], dv[NB], nrm = 0.0;
    basic_disp<D3>(f, e, U, v);
#pragma unroll
    for (int k = 0; k < NB; k++) {
      dv[k] = v[k] - vprev[k];
      nrm += dv[k] * dv[k];
      vprev[k] = v[k];
    }
    S.action = (init0 && sqrt(nrm) <= DBL_EPSILON) ? 1 : 0;
#pragma unroll
    for (int k = 0; k < NB; k++) {
      S.vin[k] = v[k] - dv[k];
      S.dvToDo[k] = dv[k];
      S.dvT[k] = dv[k];
    }
    S.converged = 0;
    S.nsub = 1;
  }
  __syncthreads();
  if (S.action == 1) return;
  const int maxSub = D3 ? 10 : 4;
  while (S.converged == 0 && S.nsub <= maxSub) {
    bool accepted = false;
    for (int l = 0; l < 3 && !accepted; l++) {
      // restart from the accepted element and section state
      if (t == 0) {
#pragma unroll
        for (int r = 0; r < NB; r++) {
          double a = 0.0;
#pragma unroll
          for (int c = 0; c < NB; c++) a += kv[r * NB + c] * S.dvT[c];
          S.SeT[r] = Se[r] + a;
        }
      }
      for (int i = w; i < nip; i += nw)
        if (lane == 0) {
          const long long sec = (long long)e * nip + i;
#pragma unroll
          for (int a = 0; a < NS; a++) {
            S.vs[i][a] = f.vs[sec * NS + a];
            S.Ssr[i][a] = f.Ssr[sec * NS + a];
          }
#pragma unroll
          for (int k = 0; k < NS * NS; k++) S.fs[i][k] = f.fsec[sec * NS * NS + k];
        }
      __syncthreads();
      const int numIters = l == 1 ? 10 * f.maxiter : f.maxiter;
      for (int j = 0; j < numIters; j++) {
        const bool initial_fs = (l == 1) || (l == 2 && j == 0);
        for (int i = w; i < nip; i += nw) force_section<D3>(f, S, e, i, lane, initial_fs, init0, L);
        __syncthreads();
        if (t == 0) {
          double fl[NB * NB], vr[NB];
#pragma unroll
          for (int k = 0; k < NB * NB; k++) fl[k] = 0.0;
#pragma unroll
          for (int k = 0; k < NB; k++) vr[k] = 0.0;
          for (int i = 0; i < nip; i++) {
#pragma unroll
            for (int k = 0; k < NB * NB; k++) fl[k] += S.fc[i][k];
#pragma unroll
            for (int k = 0; k < NB; k++) vr[k] += S.vr[i][k];
          }
          double kvT[NB * NB];
          if (!inverse_n<NB>(fl, kvT))
#pragma unroll
            for (int k = 0; k < NB * NB; k++) kvT[k] = 0.0;
          double dvr[NB], dSe[NB], dW = 0.0;
#pragma unroll
          for (int k = 0; k < NB; k++) dvr[k] = S.vin[k] + S.dvT[k] - vr[k];
#pragma unroll
          for (int r = 0; r < NB; r++) {
            double a = 0.0;
#pragma unroll
            for (int c = 0; c < NB; c++) a += kvT[r * NB + c] * dvr[c];
            dSe[r] = a;
          }
#pragma unroll
          for (int k = 0; k < NB; k++) dW += dvr[k] * dSe[k];
#pragma unroll
          for (int k = 0; k < NB; k++) S.SeT[k] += dSe[k];
          S.action = 0;
          if (fabs(dW) < f.tol) {
            double n2 = 0.0;
#pragma unroll
            for (int k = 0; k < NB; k++) {
              S.dvToDo[k] -= S.dvT[k];
              S.vin[k] += S.dvT[k];
              n2 += S.dvToDo[k] * S.dvToDo[k];
            }
            if (sqrt(n2) <= DBL_EPSILON) {
              S.converged = 1;
            } else {
#pragma unroll
              for (int k = 0; k < NB; k++) S.dvT[k] = S.dvToDo[k];
              S.nsub = 1;
            }
#pragma unroll
            for (int k = 0; k < NB * NB; k++) kv[k] = kvT[k];
#pragma unroll
            for (int k = 0; k < NB; k++) Se[k] = S.SeT[k];
            S.action = 2;
          } else if (j == numIters - 1 && l == 2) {
#pragma unroll
            for (int k = 0; k < NB; k++) S.dvT[k] /= 10.0;
            S.nsub += 1;
          }
        }
        __syncthreads();
        if (S.action == 2) {
          for (int i = w; i < nip; i += nw)
            if (lane == 0) {
              const long long sec = (long long)e * nip + i;
#pragma unroll
              for (int a = 0; a < NS; a++) {
                f.vs[sec * NS + a] = S.vs[i][a];
                f.Ssr[sec * NS + a] = S.Ssr[i][a];
              }
#pragma unroll
              for (int k = 0; k < NS * NS; k++) f.fsec[sec * NS * NS + k] = S.fs[i][k];
            }
          accepted = true;
          break;
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    if (S.converged) f.init[e] = 1;
    else err[0] = 1;
  }
}

// revertToLastCommit (ForceBeamColumn2d.cpp:310-342): section deformations
// from their committed values, fibers reverted and set to them again
// (resisting force and flexibility recomputed), basic forces and stiffness
// from their committed values, initialFlag = 0.
template <bool D3>
__global__ __launch_bounds__(kForceThreads) void k_force_revert(FrameDev f) {
  constexpr int NB = ForceShared<D3>::NB, NS = ForceShared<D3>::NS;
  const int e = blockIdx.x;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int nip = f.nip, nw = kForceThreads / 64;
  for (int i = w; i < nip; i += nw) {
    const long long sec = (long long)e * nip + i;
    section_revert<D3>(f, sec, lane);
    double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < NS; a++) d[a] = f.vs_c[sec * NS + a];
    double sr[4], kt[16], ks[NS * NS], fsn[NS * NS];
    section_fibers<D3>(f, sec, lane, d, sr, kt);
#pragma unroll
    for (int a = 0; a < NS; a++)
#pragma unroll
      for (int b = 0; b < NS; b++) ks[a * NS + b] = kt[a * (D3 ? 4 : 2) + b];
    if (!inverse_n<NS>(ks, fsn))
#pragma unroll
      for (int k = 0; k < NS * NS; k++) fsn[k] = 0.0;
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < NS; a++) {
        f.vs[sec * NS + a] = d[a];
        f.Ssr[sec * NS + a] = sr[a];
      }
#pragma unroll
      for (int k = 0; k < NS * NS; k++) f.fsec[sec * NS * NS + k] = fsn[k];
    }
  }
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < NB; k++) {
      f.Se[(long long)e * NB + k] = f.Se_c[(long long)e * NB + k];
      f.vprev[(long long)e * NB + k] = f.vprev_c[(long long)e * NB + k];
    }
#pragma unroll
    for (int k = 0; k < NB * NB; k++) f.kv[(long long)e * NB * NB + k] = f.kv_c[(long long)e * NB * NB + k];
    f.init[e] = 0;
  }
}

// Resisting force P = T^T q and tangent K = T^T kv T (getResistingForce,
// getTangentStiff; LinearCrdTransf2d/3d), one thread per element.
template <bool D3>
__global__ __launch_bounds__(64) void k_force_resid(FrameDev f, double* __restrict__ Re) {
  constexpr int NB = D3 ? 6 : 3;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= f.nele) return;
  double q[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < NB; k++) q[k] = f.Se[(long long)e * NB + k];
  global_force<D3>(f, e, q, Re);
}

template <bool D3>
__global__ __launch_bounds__(64) void k_force_tangent(FrameDev f, double* __restrict__ Ke) {
  constexpr int NB = D3 ? 6 : 3;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= f.nele) return;
  double kb[6][6];
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = 0; c < 6; c++) kb[r][c] = (r < NB && c < NB) ? f.kv[(long long)e * NB * NB + r * NB + c] : 0.0;
  global_stiff<D3>(f, e, kb, Ke);
}

void launch_force_update(hipStream_t s, const FrameDev& f, bool d3, const double* U, int* err) {
  if (f.nele <= 0) return;
  if (f.nip > kForceMaxSec) throw Error(XG_ERR_ARG, "forceBeamColumn: at most 10 sections");
  if (d3) hipLaunchKernelGGL(k_force_update<true>, dim3(f.nele), dim3(kForceThreads), 0, s, f, U, err);
  else hipLaunchKernelGGL(k_force_update<false>, dim3(f.nele), dim3(kForceThreads), 0, s, f, U, err);
}
void launch_force_revert(hipStream_t s, const FrameDev& f, bool d3) {
  if (f.nele <= 0) return;
  if (d3) hipLaunchKernelGGL(k_force_revert<true>, dim3(f.nele), dim3(kForceThreads), 0, s, f);
  else hipLaunchKernelGGL(k_force_revert<false>, dim3(f.nele), dim3(kForceThreads), 0, s, f);
}
void launch_force_resid(hipStream_t s, const FrameDev& f, bool d3, double* Re) {
  if (f.nele <= 0) return;
  const dim3 g(ceil_div(f.nele, 64)), b(64);
  if (d3) hipLaunchKernelGGL(k_force_resid<true>, g, b, 0, s, f, Re);
  else hipLaunchKernelGGL(k_force_resid<false>, g, b, 0, s, f, Re);
}
void launch_force_tangent(hipStream_t s, const FrameDev& f, bool d3, double* Ke) {
  if (f.nele <= 0) return;
  const dim3 g(ceil_div(f.nele, 64)), b(64);
  if (d3) hipLaunchKernelGGL(k_force_tangent<true>, g, b, 0, s, f, Ke);
  else hipLaunchKernelGGL(k_force_tangent<false>, g, b, 0, s, f, Ke);
}

}  // namespace xg
