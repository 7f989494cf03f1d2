// Band Cholesky of wide-banded systems (half bandwidth 64 .. kWbMaxBw), the
// exact preconditioner of the PCG for models such as RC moment frames,
// where block Jacobi needs thousands of iterations per solve.  The matrix,
// in a band order of the equations (pinv), is held as 64 x 64 tiles: tile
// column k (positions 64k .. 64k+63) keeps T + 1 tiles, tile (k, i) covering
// rows 64(k+i) .. and the columns of k, row-major; T = ceil(bw / 64), so the
// band (and the factor, which does not fill outside it) fits.  Positions
// past n pad the last tile column with identity rows.
//
// The factorisation runs one launch per tile column (k_wb_step): the
// column's workgroups apply the previous column's update to their tile,
// factor the diagonal tile (each redundantly, in LDS) and solve their tile
// against it; the remaining workgroups apply the previous column's update to
// the next T - 1 tile columns, which no workgroup of this launch reads.  The
// forward and backward sweeps run one launch per tile column as well
// (k_wb_fwd, k_wb_bwd), for one or two right-hand sides.
#include <algorithm>
#include <utility>

#include "../../include/xara_gpu.h"
#include "xg_common.h"
#include "xg_kernels.h"

namespace xg {

namespace {

constexpr int TB = 64;           // tile size
constexpr int TP = TB + 1;       // padded LDS row stride
constexpr int TS = TB * TB;      // doubles per tile

__global__ __launch_bounds__(256) void k_wb_fill(long long total, double* __restrict__ a) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) a[i] = 0.0;
}

// CSR lower triangle (in band positions) into the tiles; identity on the
// padded positions.  An entry outside the tiles' reach sets bad.
__global__ __launch_bounds__(256) void k_wb_gather(int n, int npad, int T, const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind, const double* __restrict__ vals,
                                                   const int* __restrict__ pinv, double* __restrict__ L,
                                                   int* __restrict__ bad) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= npad) return;
  if (r >= n) {
    const int k = r / TB;
    L[((long long)k * (T + 1)) * TS + (r % TB) * TB + (r % TB)] = 1.0;
    return;
  }
  const int pr = pinv[r];
  for (int q = rowptr[r]; q < rowptr[r + 1]; q++) {
    const int pc = pinv[colind[q]];
    if (pc > pr) continue;
    const int k = pc / TB, i = pr / TB - k;
    if (i > T) {
      atomicExch(bad, 2);
      continue;
    }
    L[((long long)k * (T + 1) + i) * TS + (pr % TB) * TB + (pc % TB)] = vals[q];
  }
}

// The factor and sweep kernels below run one wavefront per workgroup,
// lane r holding row r (or column r) of a tile in registers (arrays indexed
// by compile-time constants only, loops over them fully unrolled); values a
// lane needs from the other rows come from LDS (broadcast reads) or from
// readlane.  A tile row is loaded by each lane from its own 512 bytes, all
// loads issued before their first use.
__device__ __forceinline__ void row_load(double (&v)[TB], const double* __restrict__ g) {
  const double2* g2 = reinterpret_cast<const double2*>(g);
#pragma unroll
  for (int j = 0; j < TB / 2; j++) {
    const double2 d = g2[j];
    v[2 * j] = d.x;
    v[2 * j + 1] = d.y;
  }
}
__device__ __forceinline__ double lane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// The factor step runs four wavefronts per workgroup.  Thread (w, r) of
// wavefront w holds row r, columns 16w .. 16w+15 of the tile being updated
// (v16); operands are staged in LDS transposed (xt[m*TB + j] = X[j][m]) so
// that the lane's X[r][m] and row m of Y^T are plain and broadcast reads.
constexpr int WQ = TB / 4;  // columns per wavefront

// The tile GEMM W -= X Y^T: thread t forms the 4 x 4 block of rows
// 4(t/16) .., columns 4(t%16) .. (v[a][b]), from X and Y staged in LDS
// transposed: per m, four rows of column m of X and four of Y, two 16-byte
// reads each, then 16 products (a rolled loop).
__device__ __forceinline__ void block_gemm_nt_sub(double (&v)[4][4], const double* xt, const double* yt) {
  const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
#pragma unroll 2
  for (int m = 0; m < TB; m++) {
    const double2 x01 = *reinterpret_cast<const double2*>(xt + m * TB + r0);
    const double2 x23 = *reinterpret_cast<const double2*>(xt + m * TB + r0 + 2);
    const double2 y01 = *reinterpret_cast<const double2*>(yt + m * TB + c0);
    const double2 y23 = *reinterpret_cast<const double2*>(yt + m * TB + c0 + 2);
    const double xv[4] = {x01.x, x01.y, x23.x, x23.y}, yv[4] = {y01.x, y01.y, y23.x, y23.y};
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) v[a][b] -= xv[a] * yv[b];
  }
}
__device__ __forceinline__ void block_load(double (&v)[4][4], const double* __restrict__ tile) {
  const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double2* g2 = reinterpret_cast<const double2*>(tile + (r0 + a) * TB + c0);
    const double2 p0 = g2[0], p1 = g2[1];
    v[a][0] = p0.x;
    v[a][1] = p0.y;
    v[a][2] = p1.x;
    v[a][3] = p1.y;
  }
}
__device__ __forceinline__ void block_store(double* __restrict__ tile, const double (&v)[4][4]) {
  const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double2* g2 = reinterpret_cast<double2*>(tile + (r0 + a) * TB + c0);
    g2[0] = make_double2(v[a][0], v[a][1]);
    g2[1] = make_double2(v[a][2], v[a][3]);
  }
}
__device__ __forceinline__ void block_to_lds(double* wt, const double (&v)[4][4]) {
  const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) wt[(r0 + a) * TP + c0 + b] = v[a][b];
}
__device__ __forceinline__ void quarter_load(double (&v)[WQ], const double* __restrict__ tile) {
  const int r = threadIdx.x % TB, w = threadIdx.x / TB;
  const double2* g2 = reinterpret_cast<const double2*>(tile + r * TB + w * WQ);
#pragma unroll
  for (int j = 0; j < WQ / 2; j++) {
    const double2 d = g2[j];
    v[2 * j] = d.x;
    v[2 * j + 1] = d.y;
  }
}
// stage tile Y transposed: thread (w, r) loads Y[r][16w ..] and writes them
// down column r of yt (no barrier here)
__device__ __forceinline__ void stage_t(double* yt, const double* __restrict__ Y) {
  const int r = threadIdx.x % TB, w = threadIdx.x / TB;
  double y[WQ];
  quarter_load(y, Y);
#pragma unroll
  for (int m = 0; m < WQ; m++) yt[(w * WQ + m) * TB + r] = y[m];
}

// 1/sqrt(a): the hardware estimate and two Newton steps (double precision
// to a few ulp; the factor only preconditions the PCG)
__device__ __forceinline__ double rsqrt_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
  const double h = 0.5 * a;
  y = y * fma(-h, y * y, 1.5);
  y = y * fma(-h, y * y, 1.5);
  return y;
}

// Cholesky of the diagonal tile, every wavefront redundantly (lane r holding
// row r in a, lower part used); on return a[c] = L(r, c) for c <= r and rd =
// 1 / L(r, r) on lane r.  Scaled column c is broadcast through the
// wavefront's own col buffer (two, alternating): LDS accesses of one
// wavefront complete in order, so no barrier is needed, and the rank-1
// update of the later columns overlaps the next pivots.  The next pivot's
// diagonal is updated first, from the lane's own value.  One template
// instance per pivot: every register index is a compile-time constant.
template <int C>
__device__ __forceinline__ void potrf_pivot(double (&a)[TB], double& rd, double* col, bool& nonpos) {
  const int r = threadIdx.x % TB;
  double d = lane_d(a[C], C);
  const bool ok = (d > 0.0) && isfinite(d);
  nonpos |= !ok;
  d = ok ? d : 1.0;
  const double inv = rsqrt_nr(d), s = d * inv;
  const double l = (r == C) ? s : a[C] * inv;
  a[C] = l;
  if (r == C) rd = inv;
  if constexpr (C + 1 < TB) {
    if (r == C + 1) a[C + 1] -= l * l;
  }
  double* cb = col + (C & 1) * TB;
  cb[r] = l;
  // orders the compiler only: the wavefront's LDS accesses run in order
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int j = (C + 1) & ~1; j < TB; j += 2) {
    const double2 cp = *reinterpret_cast<const double2*>(cb + j);
    // lane C+1 has updated its own a[C+1] already
    if (j == C + 1) a[j] = (r == C + 1) ? a[j] : a[j] - l * cp.x;
    else if (j > C) a[j] -= l * cp.x;
    if (j + 1 == C + 1) a[j + 1] = (r == C + 1) ? a[j + 1] : a[j + 1] - l * cp.y;
    else a[j + 1] -= l * cp.y;
  }
  // keep the scheduler from hoisting later pivots' reads over this one
  // (their values would not fit in registers next to the row)
  __builtin_amdgcn_sched_barrier(0);
}
template <int... Cs>
__device__ __forceinline__ void potrf_all(std::integer_sequence<int, Cs...>, double (&a)[TB], double& rd, double* col,
                                          bool& nonpos) {
  (potrf_pivot<Cs>(a, rd, col, nonpos), ...);
}

// Column C of the right-looking solve W := W L^-T, the row's 64 entries over
// a quad of lanes (lane q holding columns 4jj + q): the owner finishes
// x_C = w_C / L(C, C), a quad broadcast hands it to the other three, and
// every lane takes x_C L(j, C) off its later columns (lkt[C*TB + j] =
// L(j, C)).
template <int C>
__device__ __forceinline__ void trsm_col(double (&w)[WQ], const double* lkt, const double* rds, int q) {
  constexpr int JJ = C / 4, QC = C % 4;
  double x = w[JJ] * rds[C];
  if (q == QC) w[JJ] = x;
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), QC * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), QC * 0x55, 0xf, 0xf, false);
  x = __hiloint2double(hi, lo);
#pragma unroll
  for (int jj = JJ; jj < WQ; jj++)
    if (4 * jj + 3 > C) {
      const double l = lkt[C * TB + 4 * jj + q];
      if (4 * jj + q > C) w[jj] -= x * l;
    }
}
template <int... Cs>
__device__ __forceinline__ void trsm_all(std::integer_sequence<int, Cs...>, double (&w)[WQ], const double* lkt,
                                         const double* rds, int q) {
  (trsm_col<Cs>(w, lkt, rds, q), ...);
}

// One tile column k of the factorisation (see the file comment).  Grid:
// T + 1 column workgroups, then T (T - 1) / 2 update workgroups; 256
// threads each.  The diagonal tiles of L go to D column-major
// (D[k][c*TB + r] = L(r, c)) for the sweeps.
__global__ __launch_bounds__(256) void k_wb_step(int k, int nt, int T, double* __restrict__ L,
                                                 double* __restrict__ D, int* __restrict__ bad) {
  __shared__ double yt[TS];        // transposed update operands
  __shared__ double xt[TS];
  __shared__ double wt[TB * TP];   // the tile in work, row-major, padded
  __shared__ double lkt[TS];       // L_kk transposed: lkt[c*TB + j] = L(j, c)
  __shared__ double col[4][2 * TB];
  __shared__ double rds[TB];
  const int b = blockIdx.x, t = threadIdx.x, r = t % TB, w = t / TB;
  auto tile = [&](int cl, int off) { return L + ((long long)cl * (T + 1) + off) * TS; };
  if (b <= T) {
    const int i = b;
    if (k + i >= nt) return;
    const bool upd = k >= 1 && T >= 1;
    {
      double v[4][4];
      block_load(v, tile(k, 0));
      if (upd) {
        // A_kk -= L(k, k-1) L(k, k-1)^T
        stage_t(yt, tile(k - 1, 1));
        __syncthreads();
        block_gemm_nt_sub(v, yt, yt);
      }
      block_to_lds(wt, v);
    }
    __syncthreads();
    // wavefront 0 factors the diagonal tile (the others would only compete
    // for the LDS broadcasts) while the others load their own tile; its rows
    // leave wt before the others reuse it
    double a[TB];
    if (w == 0) {
#pragma unroll
      for (int c = 0; c < TB; c++) a[c] = wt[r * TP + c];
    }
    __syncthreads();
    if (w == 0) {
      double rd = 1.0;
      bool nonpos = false;
      potrf_all(std::make_integer_sequence<int, TB>{}, a, rd, col[0], nonpos);
      if (nonpos && r == 0) atomicExch(bad, 1);
      if (i == 0) {
        double* Dk = D + (long long)k * TS;
#pragma unroll
        for (int c = 0; c < TB; c++) Dk[c * TB + r] = (c <= r) ? a[c] : 0.0;
      } else {
#pragma unroll
        for (int c = 0; c < TB; c++) lkt[c * TB + r] = (c <= r) ? a[c] : 0.0;
        rds[r] = rd;
      }
    }
    if (i == 0) return;
    {
      double v[4][4];
      block_load(v, tile(k, i));
      if (upd && i + 1 <= T) {
        // A(k+i, k) -= L(k+i, k-1) L(k, k-1)^T
        stage_t(xt, tile(k - 1, i + 1));
        __syncthreads();
        block_gemm_nt_sub(v, xt, yt);
      }
      block_to_lds(wt, v);
    }
    __syncthreads();
    // W := W L_kk^-T, row rr over the quad of lanes 4rr .. 4rr+3
    const int rr = t / 4, q = t % 4;
    double wv[WQ];
#pragma unroll
    for (int jj = 0; jj < WQ; jj++) wv[jj] = wt[rr * TP + 4 * jj + q];
    trsm_all(std::make_integer_sequence<int, TB>{}, wv, lkt, rds, q);
    double* dst = tile(k, i) + rr * TB;
#pragma unroll
    for (int jj = 0; jj < WQ; jj++) dst[4 * jj + q] = wv[jj];
    return;
  }
  // the previous column's update of tile columns k+1 .. k+T-1:
  // tile (k-1+a, k-1+bb) -= L(k-1+a, k-1) L(k-1+bb, k-1)^T, 2 <= bb <= a <= T
  if (k < 1) return;
  int o = b - (T + 1), bb = 2;
  while (bb <= T && o >= T - bb + 1) {
    o -= T - bb + 1;
    bb++;
  }
  if (bb > T) return;
  const int a = bb + o;
  if (k - 1 + a >= nt) return;
  double* dst = tile(k - 1 + bb, a - bb);
  double v[4][4];
  block_load(v, dst);
  stage_t(xt, tile(k - 1, a));
  stage_t(yt, tile(k - 1, bb));
  __syncthreads();
  block_gemm_nt_sub(v, xt, yt);
  block_store(dst, v);
}

// Forward sweep, tile column k: y_k = L_kk^-1 b_k on every workgroup (lane r
// holding row r of L_kk, a readlane chain over the columns), written to Yo
// by workgroup 0; workgroup i >= 1: b_{k+i} -= L(k+i, k) y_k.  R right-hand
// sides, each npad long.
template <int R>
__global__ __launch_bounds__(64) void k_wb_fwd(int k, int nt, int T, int npad, const double* __restrict__ L,
                                               const double* __restrict__ D, double* __restrict__ Yb,
                                               double* __restrict__ Yo) {
  const int i = blockIdx.x, r = threadIdx.x;
  if (k + i >= nt) return;
  const double* Dk = D + (long long)k * TS;
  double l[TB], t[TB];
#pragma unroll
  for (int c = 0; c < TB; c++) l[c] = Dk[c * TB + r];
  if (i >= 1) row_load(t, L + ((long long)k * (T + 1) + i) * TS + r * TB);
  double bv[R], y[R];
#pragma unroll
  for (int s = 0; s < R; s++) {
    bv[s] = Yb[(long long)s * npad + k * TB + r];
    y[s] = 0.0;
  }
  double dg = 1.0;
#pragma unroll
  for (int c = 0; c < TB; c++)
    if (r == c) dg = l[c];
  const double rdl = 1.0 / dg;
#pragma unroll
  for (int c = 0; c < TB; c++) {
    const double rc = lane_d(rdl, c);
#pragma unroll
    for (int s = 0; s < R; s++) {
      const double yc = lane_d(bv[s], c) * rc;
      if (r == c) y[s] = yc;
      bv[s] -= l[c] * yc;
    }
  }
  if (i == 0) {
#pragma unroll
    for (int s = 0; s < R; s++) Yo[(long long)s * npad + k * TB + r] = y[s];
    return;
  }
#pragma unroll
  for (int s = 0; s < R; s++) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < TB; j++) acc += t[j] * lane_d(y[s], j);
    Yb[(long long)s * npad + (k + i) * TB + r] -= acc;
  }
}

// Backward sweep, tile column k (k from the last down): x_k = L_kk^-T z_k
// on every workgroup (lane r holding column r of L_kk), written to Xo by
// workgroup 0; workgroup j >= 1: z_{k-j} -= L(k, k-j)^T x_k, z kept in Yo.
template <int R>
__global__ __launch_bounds__(64) void k_wb_bwd(int k, int nt, int T, int npad, const double* __restrict__ L,
                                               const double* __restrict__ D, double* __restrict__ Yo,
                                               double* __restrict__ Xo) {
  const int j = blockIdx.x, r = threadIdx.x;
  if (k - j < 0) return;
  const double* Dk = D + (long long)k * TS;
  double l[TB], t[TB];
  row_load(l, Dk + r * TB);  // column r of L_kk: l[c] = L(c, r)
  if (j >= 1) {
    // tile (k-j, j) holds L(k, k-j), row-major: lane r takes its column r
    const double* Lt = L + ((long long)(k - j) * (T + 1) + j) * TS;
#pragma unroll
    for (int q = 0; q < TB; q++) t[q] = Lt[q * TB + r];
  }
  double zv[R], x[R];
#pragma unroll
  for (int s = 0; s < R; s++) {
    zv[s] = Yo[(long long)s * npad + k * TB + r];
    x[s] = 0.0;
  }
  double dg = 1.0;
#pragma unroll
  for (int c = 0; c < TB; c++)
    if (r == c) dg = l[c];
  const double rdl = 1.0 / dg;
#pragma unroll
  for (int c = TB - 1; c >= 0; c--) {
    const double rc = lane_d(rdl, c);
#pragma unroll
    for (int s = 0; s < R; s++) {
      const double xc = lane_d(zv[s], c) * rc;
      if (r == c) x[s] = xc;
      zv[s] -= l[c] * xc;
    }
  }
  if (j == 0) {
#pragma unroll
    for (int s = 0; s < R; s++) Xo[(long long)s * npad + k * TB + r] = x[s];
    return;
  }
#pragma unroll
  for (int s = 0; s < R; s++) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < TB; q++) acc += t[q] * lane_d(x[s], q);
    Yo[(long long)s * npad + (k - j) * TB + r] -= acc;
  }
}

// to the band order (padding zeroed) and back
__global__ __launch_bounds__(256) void k_wb_perm_in(int n, int npad, const int* __restrict__ pinv,
                                                    const double* __restrict__ v, double* __restrict__ y) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < n) y[pinv[p]] = v[p];
  else if (p < npad) y[p] = 0.0;
}
__global__ __launch_bounds__(256) void k_wb_perm_out(int n, const int* __restrict__ pinv,
                                                     const double* __restrict__ y, double* __restrict__ v) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < n) v[p] = y[pinv[p]];
}

}  // namespace

int wband_tiles(int bw) { return (bw + TB - 1) / TB; }
int wband_padded(int n) { return ((n + TB - 1) / TB) * TB; }
void wband_zero(hipStream_t s, long long n, double* v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_wb_fill, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, s, n, v);
}
size_t wband_factor_doubles(int n, int bw) {
  const long long nt = (n + TB - 1) / TB;
  return (size_t)(nt * (wband_tiles(bw) + 1) * TS);
}
size_t wband_diag_doubles(int n) { return (size_t)((n + TB - 1) / TB) * TS; }

void wband_factor(hipStream_t s, int n, int bw, const int* rowptr, const int* colind, const double* vals,
                  const int* pinv, double* L, double* D, int* bad) {
  if (n <= 0) return;
  const int nt = (n + TB - 1) / TB, T = wband_tiles(bw), npad = nt * TB;
  const long long total = (long long)nt * (T + 1) * TS;
  hipLaunchKernelGGL(k_wb_fill, dim3((unsigned)std::min<long long>((total + 255) / 256, 8192)), dim3(256), 0, s,
                     total, L);
  hipLaunchKernelGGL(k_wb_gather, dim3((npad + 255) / 256), dim3(256), 0, s, n, npad, T, rowptr, colind, vals, pinv,
                     L, bad);
  const int grid = (T + 1) + T * (T - 1) / 2;
  for (int k = 0; k < nt; k++)
    hipLaunchKernelGGL(k_wb_step, dim3(grid), dim3(256), 0, s, k, nt, T, L, D, bad);
}

void wband_perm_in(hipStream_t s, int n, int bw, const int* pinv, const double* v, double* y) {
  const int npad = ((n + TB - 1) / TB) * TB;
  hipLaunchKernelGGL(k_wb_perm_in, dim3((npad + 255) / 256), dim3(256), 0, s, n, npad, pinv, v, y);
  (void)bw;
}
void wband_perm_out(hipStream_t s, int n, const int* pinv, const double* y, double* v) {
  hipLaunchKernelGGL(k_wb_perm_out, dim3((n + 255) / 256), dim3(256), 0, s, n, pinv, y, v);
}

void wband_sweeps(hipStream_t s, int n, int bw, int nrhs, const double* L, const double* D, double* Yb, double* Yo,
                  double* Xo) {
  if (n <= 0) return;
  const int nt = (n + TB - 1) / TB, T = wband_tiles(bw), npad = nt * TB;
  if (nrhs != 1 && nrhs != 2) throw Error(XG_ERR_ARG, "band sweeps: one or two right-hand sides");
  for (int k = 0; k < nt; k++) {
    if (nrhs == 1) hipLaunchKernelGGL(k_wb_fwd<1>, dim3(T + 1), dim3(64), 0, s, k, nt, T, npad, L, D, Yb, Yo);
    else hipLaunchKernelGGL(k_wb_fwd<2>, dim3(T + 1), dim3(64), 0, s, k, nt, T, npad, L, D, Yb, Yo);
  }
  for (int k = nt - 1; k >= 0; k--) {
    if (nrhs == 1) hipLaunchKernelGGL(k_wb_bwd<1>, dim3(T + 1), dim3(64), 0, s, k, nt, T, npad, L, D, Yo, Xo);
    else hipLaunchKernelGGL(k_wb_bwd<2>, dim3(T + 1), dim3(64), 0, s, k, nt, T, npad, L, D, Yo, Xo);
  }
}

}  // namespace xg
