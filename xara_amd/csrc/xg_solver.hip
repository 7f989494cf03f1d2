// Assembly into the fixed-pattern CSR system and the preconditioned CG
// LinearSOE on gfx950.
//
// Assembly is a gather: every CSR slot (and every right-hand-side entry)
// owns the list of element-matrix entries that land on it, stored in
// element-tag order, so each value is summed in the order of
// IncrementalIntegrator::formTangent / formElementResidual
// (IncrementalIntegrator.cpp:74-98, :221-236) with no floating-point atomics.
//
// Dot products reduce per block with a fixed LDS tree and then over blocks
// in a single-block pass, so every PCG iteration is bitwise reproducible.
#include <algorithm>
#include "../../include/xara_gpu.h"
#include "xg_common.h"
#include "xg_kernels.h"

namespace xg {

// ---------------------------------------------------------------------------
// Assembly
// ---------------------------------------------------------------------------
__global__ void k_assemble_csr(long long nnz, const int* __restrict__ ptr, const int* __restrict__ idx,
                               const double* __restrict__ Ke, double* __restrict__ vals) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double acc = 0.0;
  const int b = ptr[k], e = ptr[k + 1];
  for (int j = b; j < e; j++) acc += Ke[idx[j]];
  vals[k] = acc;
}

// B[a] = sum over elements (in tag order) of -R_e + lambda * P[a]
__global__ void k_assemble_rhs(int neq, const int* __restrict__ ptr, const int* __restrict__ idx,
                               const double* __restrict__ Re, double lambda,
                               const double* __restrict__ P, double* __restrict__ B) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= neq) return;
  double acc = 0.0;
  for (int j = ptr[a]; j < ptr[a + 1]; j++) acc += -Re[idx[j]];
  if (P) acc += lambda * P[a];
  B[a] = acc;
}

// Transient right-hand side (TransientIntegrator::formUnbalance with
// Node::getUnbalancedLoadIncInertia, Node.cpp:740-763, and the
// EarthquakePattern inertia load -M R a_g, Node.cpp:668-688):
//   B[a] = sum_e -R_e + (lambda P[a] - m r a_g - m A - alphaM m V)
__global__ void k_assemble_rhs_dyn(int neq, const int* __restrict__ ptr, const int* __restrict__ idx,
                                   const double* __restrict__ Re, double lambda, const double* __restrict__ P,
                                   const double* __restrict__ Meq, const double* __restrict__ Req, double ag,
                                   double alphaM, const int* __restrict__ eq2dof, const double* __restrict__ V,
                                   const double* __restrict__ A, double* __restrict__ B) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= neq) return;
  double acc = 0.0;
  for (int j = ptr[a]; j < ptr[a + 1]; j++) acc += -Re[idx[j]];
  double u = 0.0;
  if (P) u = lambda * P[a];
  const double m = Meq[a];
  if (m != 0.0) {
    const long long d = eq2dof[a];
    u += -(m * Req[a]) * ag;
    u += -m * A[d];
    if (alphaM != 0.0) u += -alphaM * m * V[d];
  }
  B[a] = acc + u;
}

// Newmark predictor on free DOFs (Newmark::newStep, Newmark.cpp:105-170):
// U = U_n, V = a1 V_n + a2 A_n, A = a4 A_n + a3 V_n
__global__ void k_newmark_predict(long long ndofs, const int* __restrict__ eqmap, const double* __restrict__ Uc,
                                  const double* __restrict__ Vc, const double* __restrict__ Ac, double a1,
                                  double a2, double a3, double a4, double* __restrict__ U, double* __restrict__ V,
                                  double* __restrict__ A) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndofs || eqmap[i] < 0) return;
  U[i] = Uc[i];
  V[i] = Vc[i] * a1 + Ac[i] * a2;
  A[i] = Ac[i] * a4 + Vc[i] * a3;
}

// Newmark::update (Newmark.cpp:411-458): U += dU, V += c2 dU, A += c3 dU
__global__ void k_newmark_correct(long long ndofs, const int* __restrict__ eqmap, const double* __restrict__ dU,
                                  double c2, double c3, double* __restrict__ U, double* __restrict__ V,
                                  double* __restrict__ A) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndofs) return;
  const int a = eqmap[i];
  if (a < 0) return;
  const double d = dU[a];
  U[i] += d;
  V[i] += c2 * d;
  A[i] += c3 * d;
}

// Diagonal add (nodal masses etc.): vals[pos[i]] += coef * m[i]
__global__ void k_add_diag(int n, const long long* __restrict__ pos, const double* __restrict__ m,
                           double coef, double* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vals[pos[i]] += coef * m[i];
}

// ---------------------------------------------------------------------------
// Block reductions (deterministic: fixed butterfly and wave order)
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
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

// ---------------------------------------------------------------------------
// PCG scalar state kept on the device
// ---------------------------------------------------------------------------
// r.z and the iteration count are kept per iteration parity: iteration k
// reads entry k & 1 and writes entry (k + 1) & 1, so that the workgroups of
// a fused kernel that all recompute the same reduction never read a field
// that one of them rewrites.
struct CGScal {
  double rzp[2], pq, rr, bb, alpha, beta, tol2;
  int converged, iterp[2], maxit, breakdown;
  int fresh;  // no update yet in this solve (the first new direction is z itself)
};

// Sum of `nb` partials by one workgroup of 256: a thread's partials in
// ascending order (up to 32 loads in flight, so 8,192 partials cost one
// memory round trip), then the tree sum.  Every caller sums in this order,
// so every workgroup that reduces the same partials gets the same bits.
__device__ __forceinline__ double sum_partials(int nb, const double* __restrict__ part, double* sh) {
  double v = 0.0;
  int i = threadIdx.x;
  for (; i + 31 * 256 < nb; i += 32 * 256) {
    double a[32];
#pragma unroll
    for (int u = 0; u < 32; u++) a[u] = part[i + u * 256];
#pragma unroll
    for (int u = 0; u < 32; u++) v += a[u];
  }
  for (; i + 7 * 256 < nb; i += 8 * 256) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = part[i + u * 256];
#pragma unroll
    for (int u = 0; u < 8; u++) v += a[u];
  }
  for (; i < nb; i += 256) v += part[i];
  return block_sum<256>(v, sh);
}

// Sums of the `nb` interleaved partial pairs (r.z, r.r) of the update, each
// thread's pairs in ascending order with eight pairs' loads in flight.
__device__ __forceinline__ void sum_partial_pairs(int nb, const double* __restrict__ part, double* sh, double& a,
                                                  double& b) {
  double v1 = 0.0, v2 = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * 256 < nb; i += 8 * 256) {
    double c[8], d[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      c[u] = part[2 * (i + u * 256)];
      d[u] = part[2 * (i + u * 256) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      v1 += c[u];
      v2 += d[u];
    }
  }
  for (; i < nb; i += 256) {
    v1 += part[2 * i];
    v2 += part[2 * i + 1];
  }
  a = block_sum<256>(v1, sh);
  b = block_sum<256>(v2, sh);
}

// The converged flag as one value for the whole workgroup.  In a fused
// kernel workgroup 0 may set the flag while another workgroup's waves are
// still reading it; one read by thread 0 keeps its waves on one path (and
// its block sums whole).
__device__ __forceinline__ bool block_converged(const CGScal* s, int* flag) {
  if (threadIdx.x == 0) *flag = s->converged;
  __syncthreads();
  return *flag != 0;
}

// On-the-fly new direction for the fused small-system PCG: the product of
// iteration k reads p = z + beta p_old for its columns (beta from the
// partials of iteration k-1's update, reduced by every workgroup in the
// order of k_cg_beta, with the convergence test of k_cg_beta) and writes p
// of its own rows to the other buffer; the separate new-direction kernel
// goes away.  The first product of a solve (fresh) takes beta = 0, p = z.
struct OtfArgs {
  const double* z;
  const double* pold;
  double* pnew;
  CGScal* S;
  int par;      // parity of the iteration of the product
  int nbv;      // partial pairs of the previous update
  const double* part2;
};

// beta and the convergence test of the previous iteration; false when the
// solve has stopped (the product is skipped).  Workgroup-uniform.
__device__ __forceinline__ bool otf_beta(const OtfArgs& o, double* sh, int* flag, double& beta) {
  if (block_converged(o.S, flag)) return false;
  if (o.S->fresh) {
    beta = 0.0;
    return true;
  }
  double rz, rr;
  sum_partial_pairs(o.nbv, o.part2, sh, rz, rr);
  const int bpar = o.par ^ 1;  // the previous iteration's parity
  const int iter = o.S->iterp[bpar] + 1;
  beta = rz / o.S->rzp[bpar];
  const bool conv = rr <= o.S->tol2 * o.S->bb || iter >= o.S->maxit;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    o.S->iterp[o.par] = iter;
    o.S->rr = rr;
    o.S->beta = beta;
    o.S->rzp[o.par] = rz;
    if (conv) o.S->converged = 1;
  }
  return !conv;
}

// ---------------------------------------------------------------------------
// CSR SpMV y = A x with L lanes per row; per-block partial of x.y
// ---------------------------------------------------------------------------
// One lane's share of a row: sum over k = b + lane, b + lane + L, ... < e of
// vals[k] * x[col(raw(k), k)], in that order.  raw(k) is the stored column
// index of entry k and col() forms the column from it.  The value and index
// loads of U entries are issued together, then their x loads, then the
// products, so a wave waits for memory twice per U entries rather than twice
// per entry.  Entries past e load entry e - 1 again (in bounds, cached) and
// add nothing.
template <int L, int U = 4, class XR, class Raw, class Col>
__device__ __forceinline__ double row_dot(int b, int e, int lane, const double* __restrict__ vals, XR xr, Raw raw,
                                          Col col) {
  double sum = 0.0;
  for (int k0 = b + lane; k0 < e; k0 += U * L) {
    double v[U], xv[U];
    int r[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = min(k0 + u * L, e - 1);
      v[u] = vals[k];
      r[u] = raw(k);
    }
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = xr(col(r[u], min(k0 + u * L, e - 1)));
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u * L < e) sum += v[u] * xv[u];
  }
  return sum;
}

template <int L, bool OTF = false>
__global__ __launch_bounds__(256) void k_spmv_dot(int n, const int* __restrict__ rowptr,
                                                  const int* __restrict__ colind,
                                                  const double* __restrict__ vals,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ partial,
                                                  const int* __restrict__ stop, OtfArgs o) {
  __shared__ double sh[256];
  __shared__ int cflag;
  double beta = 0.0;
  if constexpr (OTF) {
    if (!otf_beta(o, sh, &cflag, beta)) return;
  } else {
    if (stop && stop[0]) return;
  }
  auto xr = [&](int c) -> double {
    if constexpr (OTF) return o.z[c] + beta * o.pold[c];
    else return x[c];
  };
  const int t = threadIdx.x;
  const int lane = t % L;
  constexpr int R = 256 / L;  // rows per block and sweep
  double d = 0.0;
  // grid-stride over row groups: a fixed assignment, so the partial of each
  // block (and the dot product) does not depend on timing
  for (int row0 = blockIdx.x * R; row0 < n; row0 += gridDim.x * R) {
    const int row = row0 + t / L;
    double sum = 0.0;
    if (row < n) {
      const int b = rowptr[row], e = rowptr[row + 1];
      sum = row_dot<L>(
          b, e, lane, vals, xr, [&](int k) { return colind[k]; }, [](int c, int) { return c; });
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_down(sum, off, L);
    if (row < n && lane == 0) {
      y[row] = sum;
      const double xrow = xr(row);
      if constexpr (OTF) o.pnew[row] = xrow;
      d += sum * xrow;
    }
  }
  if (partial) {
    const double bs = block_sum<256>(d, sh);
    if (t == 0) partial[blockIdx.x] = bs;
  }
}

// ---------------------------------------------------------------------------
// Node-blocked SpMV.  When every node owns M consecutive equations (all its
// DOFs free, M = ndf) or none, equation M*J + r belongs to node block J and
// its CSR row lists the neighbouring blocks' equations as ascending runs of
// M, the same runs in each of the M rows of a block.  The columns then follow
// from one block column per neighbour (nbcol, 4 bytes per M*M nonzeros
// instead of 4 per nonzero) and rowptr[M*J] / (M*M) is block J's first entry
// in nbcol; the values are the CSR values themselves.  Same lane layout,
// reductions and dot-product partials as k_spmv_dot.
// ---------------------------------------------------------------------------
template <int L, int M, bool OTF = false>
__global__ __launch_bounds__(256) void k_spmv_nb_dot(int n, const int* __restrict__ rowptr,
                                                     const int* __restrict__ nbcol,
                                                     const double* __restrict__ vals,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     double* __restrict__ partial,
                                                     const int* __restrict__ stop, OtfArgs o) {
  __shared__ double sh[256];
  __shared__ int cflag;
  double beta = 0.0;
  if constexpr (OTF) {
    if (!otf_beta(o, sh, &cflag, beta)) return;
  } else {
    if (stop && stop[0]) return;
  }
  auto xr = [&](int c) -> double {
    if constexpr (OTF) return o.z[c] + beta * o.pold[c];
    else return x[c];
  };
  const int t = threadIdx.x;
  const int lane = t % L;
  constexpr int R = 256 / L;
  double d = 0.0;
  for (int row0 = blockIdx.x * R; row0 < n; row0 += gridDim.x * R) {
    const int row = row0 + t / L;
    double sum = 0.0;
    if (row < n) {
      const int b = rowptr[row], e = rowptr[row + 1];
      // the M rows of a block have the same length, so the block's first
      // entry is b - (row % M) * (e - b)
      const int* nc = nbcol + (b - (row % M) * (e - b)) / (M * M);
      sum = row_dot<L>(
          b, e, lane, vals, xr, [&](int k) { return nc[(k - b) / M]; },
          [&](int c, int k) { return c * M + (k - b) % M; });
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_down(sum, off, L);
    if (row < n && lane == 0) {
      y[row] = sum;
      const double xrow = xr(row);
      if constexpr (OTF) o.pnew[row] = xrow;
      d += sum * xrow;
    }
  }
  if (partial) {
    const double bs = block_sum<256>(d, sh);
    if (t == 0) partial[blockIdx.x] = bs;
  }
}

// ---------------------------------------------------------------------------
// Block-symmetric SpMV.  The consistent tangents of this path are symmetric,
// so a node-blocked matrix is kept a second time as its upper M x M blocks
// only (block-major, row-major inside a block, copied from the CSR values
// after each assembly): about half the bytes per product.  Block row I lists
// its neighbours J ascending, each with the stored block that holds it: its
// own (I, J), or its partner's (J, I) used transposed (sign bit of blk).  The
// block rows are swept in the layout's order (the equation order, or the
// Morton order of the nodes, the vectors then in that order too).  A stored
// block is read twice per product, by its own row and by its partner's, so
// the second read may find it in a cache.  Units of 256/G block rows are
// dealt out to the XCDs in contiguous eighths (the hardware dispatches
// workgroup b to XCD b % 8), one unit per workgroup (up to kSymMaxBlocks
// units, then contiguous runs): the workgroups resident on an XCD then sweep
// one front of its rows, and a row's partners, read shortly before by the
// same XCD, are more often found in its L2 (about 8% faster at 64^3 than
// contiguous runs of units per workgroup, 12% than a grid-stride sweep).  G
// lanes per block row, one neighbour block per lane and pass (its M*M values,
// column and M entries of x loaded together), the M row sums reduced across
// the G lanes.  The kernel is bound by the L1's address processing (TA and
// TD 80-84% busy at 64^3, about 33 cache lines per load instruction: each
// lane's 72-byte block arrives in five 16-byte pieces from lines of its own),
// not by HBM.  (Tried and slower: staging a row's blocks in LDS with
// consecutive lanes reading consecutive doubles, 1.7x; staging a wavefront's
// own blocks in LDS by coalesced loads behind a barrier, 10%; two
// neighbours per lane in flight at once, 5%.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void xcd_units(int units, int& ua, int& ub) {
  const int nx = min(8, (int)gridDim.x);
  const int xcd = blockIdx.x % nx, l = blockIdx.x / nx;
  const int gx = ((int)gridDim.x - xcd + nx - 1) / nx;  // workgroups of this XCD
  const int u0 = (int)((long long)units * xcd / nx), u1 = (int)((long long)units * (xcd + 1) / nx);
  const int cnt = (u1 - u0 + gx - 1) / gx;
  ua = min(u1, u0 + l * cnt);
  ub = min(u1, ua + cnt);
}

template <int M, int G, bool OTF = false>
__global__ __launch_bounds__(256) void k_spmv_sym_dot(int nbr, const int* __restrict__ ptr,
                                                      const int* __restrict__ col, const int* __restrict__ blk,
                                                      const double* __restrict__ sv, const double* __restrict__ x,
                                                      double* __restrict__ y, double* __restrict__ partial,
                                                      const int* __restrict__ stop, OtfArgs o) {
  constexpr int R = 256 / G;  // block rows per workgroup and unit
  constexpr int MM = M * M;
  __shared__ double sh[256];
  __shared__ int cflag;
  double beta = 0.0;
  if constexpr (OTF) {
    if (!otf_beta(o, sh, &cflag, beta)) return;
  } else {
    if (stop && stop[0]) return;
  }
  auto xr = [&](int c) -> double {
    if constexpr (OTF) return o.z[c] + beta * o.pold[c];
    else return x[c];
  };
  const int t = threadIdx.x;
  const int lane = t % G;
  double d = 0.0;
  int ua, ub;
  xcd_units((nbr + R - 1) / R, ua, ub);
  for (int u = ua; u < ub; u++) {
    const int I = u * R + t / G;
    double acc[M];
#pragma unroll
    for (int r = 0; r < M; r++) acc[r] = 0.0;
    if (I < nbr) {
      const int b = ptr[I], e = ptr[I + 1];
      for (int k = b + lane; k < e; k += G) {
        const int J = col[k];
        const int bk = blk[k];
        const double* a = sv + (long long)(bk & 0x7fffffff) * MM;
        double v[MM], xj[M];
#pragma unroll
        for (int q = 0; q < MM; q++) v[q] = a[q];
#pragma unroll
        for (int c = 0; c < M; c++) xj[c] = xr(M * J + c);
        const bool lower = bk < 0;  // the partner's stored block, transposed
#pragma unroll
        for (int r = 0; r < M; r++)
#pragma unroll
          for (int c = 0; c < M; c++) acc[r] += (lower ? v[c * M + r] : v[r * M + c]) * xj[c];
      }
    }
#pragma unroll
    for (int r = 0; r < M; r++)
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) acc[r] += __shfl_down(acc[r], off, G);
    if (I < nbr && lane == 0) {
#pragma unroll
      for (int r = 0; r < M; r++) {
        y[M * I + r] = acc[r];
        const double xrow = xr(M * I + r);
        if constexpr (OTF) o.pnew[M * I + r] = xrow;
        d += acc[r] * xrow;
      }
    }
  }
  if (partial) {
    const double bs = block_sum<256>(d, sh);
    if (t == 0) partial[blockIdx.x] = bs;
  }
}

// Stored block k = rows M*I.. of the CSR values from entry src[k] of row M*I
// (the block's first column), len[k] entries per row of block row I.
template <int M>
__global__ __launch_bounds__(256) void k_sym_convert(int nup, const int* __restrict__ src,
                                                     const int* __restrict__ len, const double* __restrict__ vals,
                                                     double* __restrict__ sv) {
  for (int k = blockIdx.x * 256 + threadIdx.x; k < nup; k += gridDim.x * 256) {
    const long long s0 = src[k];
    const int L = len[k];
    double v[M * M];
#pragma unroll
    for (int r = 0; r < M; r++)
#pragma unroll
      for (int c = 0; c < M; c++) v[r * M + c] = vals[s0 + (long long)r * L + c];
#pragma unroll
    for (int q = 0; q < M * M; q++) sv[(long long)k * (M * M) + q] = v[q];
  }
}

// Rows of B doubles between the equation order and the sweep order of the
// symmetric product: dst[p*B + t] = src[ord[p]*B + t] (to the sweep order),
// or dst[ord[p]*B + t] = src[p*B + t] (back).
__global__ __launch_bounds__(256) void k_sweep_perm(int nbr, int B, const int* __restrict__ ord,
                                                    const double* __restrict__ src, double* __restrict__ dst,
                                                    int back) {
  const long long total = (long long)nbr * B;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const long long p = i / B, t = i - p * B;
    const long long o = (long long)ord[p] * B + t;
    if (back) dst[o] = src[i];
    else dst[i] = src[o];
  }
}

// alpha = r.z / p.q from the SpMV's partials of p.q (one workgroup).
__global__ __launch_bounds__(256) void k_cg_alpha(int nb, const double* __restrict__ part,
                                                  CGScal* __restrict__ s, int par) {
  __shared__ double sh[256];
  if (s->converged) return;
  const double pq = sum_partials(nb, part, sh);
  if (threadIdx.x == 0) {
    s->pq = pq;
    if (pq == 0.0 || !isfinite(pq)) {
      s->breakdown = 1;
      s->converged = 1;
    } else {
      s->alpha = s->rzp[par] / pq;
    }
  }
}

// x += alpha p; r_out = r_in - alpha q; z = M^{-1} r_out; partials of r.z
// and r.r.  M^{-1}: block rows binv[i*BW + j] over the nodal block starting
// at bfirst[i] with bsize[i] <= BW entries (BW = 1 gives Jacobi).  r is double
// buffered, so a row's block neighbours are recomputed from r_in with the
// same arithmetic instead of being read while another thread block writes.
// The BW loads of a row are issued together (clamped into the block) and the
// products summed in column order, so a small system waits for memory once
// per row rather than once per block column.
template <int BW>
__device__ __forceinline__ void cg_update_rows(int n, double alpha, const double* __restrict__ p,
                                               const double* __restrict__ q, double* __restrict__ x,
                                               const double* __restrict__ r_in, double* __restrict__ r_out,
                                               double* __restrict__ z, const double* __restrict__ binv,
                                               const int* __restrict__ bfirst,
                                               const unsigned char* __restrict__ bsize, double* __restrict__ part,
                                               double* sh) {
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    x[i] += alpha * p[i];
    const double ri = r_in[i] - alpha * q[i];
    r_out[i] = ri;
    double zi;
    if constexpr (BW == 1) {
      zi = binv[i] * ri;
    } else {
      const int f = bfirst[i], m = bsize[i];
      double w[BW], rc[BW];
#pragma unroll
      for (int j = 0; j < BW; j++) {
        const int c = f + min(j, m - 1);
        w[j] = binv[(long long)i * BW + j];
        rc[j] = r_in[c] - alpha * q[c];
      }
      zi = 0.0;
#pragma unroll
      for (int j = 0; j < BW; j++)
        if (j < m) zi += w[j] * rc[j];
    }
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  const double s1 = block_sum<256>(rz, sh);
  const double s2 = block_sum<256>(rr, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

template <int BW>
__global__ __launch_bounds__(256) void k_cg_update(int n, const CGScal* __restrict__ s,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   const double* __restrict__ r_in,
                                                   double* __restrict__ r_out, double* __restrict__ z,
                                                   const double* __restrict__ binv,
                                                   const int* __restrict__ bfirst,
                                                   const unsigned char* __restrict__ bsize,
                                                   double* __restrict__ part) {
  __shared__ double sh[256];
  if (s->converged) return;
  cg_update_rows<BW>(n, s->alpha, p, q, x, r_in, r_out, z, binv, bfirst, bsize, part, sh);
}

// k_cg_alpha fused into the update, for systems whose SpMV leaves few
// partials: every workgroup sums the nbs partials of p.q itself (same order,
// same bits as k_cg_alpha) and workgroup 0 records pq and alpha.  The
// update's partials go to part_out, apart from the SpMV's.
template <int BW>
__global__ __launch_bounds__(256) void k_cg_update_fused(int n, CGScal* __restrict__ s, int par, int nbs,
                                                         const double* __restrict__ part_in,
                                                         const double* __restrict__ p,
                                                         const double* __restrict__ q, double* __restrict__ x,
                                                         const double* __restrict__ r_in,
                                                         double* __restrict__ r_out, double* __restrict__ z,
                                                         const double* __restrict__ binv,
                                                         const int* __restrict__ bfirst,
                                                         const unsigned char* __restrict__ bsize,
                                                         double* __restrict__ part_out) {
  __shared__ double sh[256];
  __shared__ int cflag;
  if (block_converged(s, &cflag)) return;
  const double pq = sum_partials(nbs, part_in, sh);
  const bool bad = pq == 0.0 || !isfinite(pq);
  const double alpha = bad ? 0.0 : s->rzp[par] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->pq = pq;
    s->fresh = 0;
    if (bad) {
      s->breakdown = 1;
      s->converged = 1;
    } else {
      s->alpha = alpha;
    }
  }
  if (bad) return;
  cg_update_rows<BW>(n, alpha, p, q, x, r_in, r_out, z, binv, bfirst, bsize, part_out, sh);
}

// beta = r.z (new) / r.z (old) and the convergence test (one workgroup).
__global__ __launch_bounds__(256) void k_cg_beta(int nb, const double* __restrict__ part,
                                                 CGScal* __restrict__ s, int par) {
  __shared__ double sh[256];
  if (s->converged) return;
  double rz, rr;
  sum_partial_pairs(nb, part, sh, rz, rr);
  if (threadIdx.x == 0) {
    const int iter = s->iterp[par] + 1;
    s->iterp[par ^ 1] = iter;
    s->rr = rr;
    s->beta = rz / s->rzp[par];
    s->rzp[par ^ 1] = rz;
    if (rr <= s->tol2 * s->bb || iter >= s->maxit) s->converged = 1;
  }
}

__global__ void k_cg_newp(int n, const CGScal* __restrict__ s, const double* __restrict__ z,
                          double* __restrict__ p) {
  if (s->converged) return;
  const double beta = s->beta;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

// Start: x = 0, r = b, z = M^{-1} b, p = z; partials of r.z, b.b
// x = 0, r = b (q == nullptr), or, from the initial guess already in x and
// q = A x, r = b - q; z = M^-1 r, p = z, and per workgroup the partials of
// r.z, b.b and r.r.
__global__ __launch_bounds__(256) void k_cg_start(int n, const double* __restrict__ b,
                                                  const double* __restrict__ q, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ z,
                                                  double* __restrict__ p, int bw, const double* __restrict__ binv,
                                                  const int* __restrict__ bfirst,
                                                  const unsigned char* __restrict__ bsize,
                                                  double* __restrict__ part) {
  __shared__ double sh[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double rz = 0.0, bb = 0.0, rr = 0.0;
  if (i < n) {
    auto res = [&](int k) { return q ? b[k] - q[k] : b[k]; };
    if (!q) x[i] = 0.0;
    const double ri = res(i);
    r[i] = ri;
    double zi;
    if (bw == 1) zi = binv[i] * ri;
    else {
      zi = 0.0;
      const int f = bfirst[i], m = bsize[i];
      for (int j = 0; j < m; j++) zi += binv[(long long)i * bw + j] * res(f + j);
    }
    z[i] = zi;
    p[i] = zi;
    rz = ri * zi;
    bb = b[i] * b[i];
    rr = ri * ri;
  }
  const double s1 = block_sum<256>(rz, sh);
  const double s2 = block_sum<256>(bb, sh);
  const double s3 = block_sum<256>(rr, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s1;
    part[3 * blockIdx.x + 1] = s2;
    part[3 * blockIdx.x + 2] = s3;
  }
}

__global__ __launch_bounds__(256) void k_cg_init_scal(int nb, const double* __restrict__ part,
                                                      double tol2, int maxit, CGScal* __restrict__ s) {
  __shared__ double sh[256];
  double v1 = 0.0, v2 = 0.0, v3 = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    v1 += part[3 * i];
    v2 += part[3 * i + 1];
    v3 += part[3 * i + 2];
  }
  const double rz = block_sum<256>(v1, sh);
  const double bb = block_sum<256>(v2, sh);
  const double rr = block_sum<256>(v3, sh);
  if (threadIdx.x == 0) {
    s->rzp[0] = rz;
    s->rzp[1] = 0.0;
    s->bb = bb;
    s->rr = rr;
    s->tol2 = tol2;
    s->maxit = maxit;
    s->iterp[0] = s->iterp[1] = 0;
    s->fresh = 1;
    s->breakdown = 0;
    s->alpha = s->beta = s->pq = 0.0;
    // an initial guess may already meet the tolerance
    s->converged = (bb == 0.0 || rr <= tol2 * bb) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Whole PCG solve of a small system in one workgroup of 1024 threads: the
// same iteration as cg_start + cg_iterations (x = 0, r = b, z = M^{-1} r,
// alpha = r.z / p.q, beta = r.z (new) / r.z (old), stop when r.r <= tol2 b.b
// or after maxit iterations, breakdown on a zero or non-finite p.q), with
// workgroup barriers in place of kernel boundaries and no host round trip
// per chunk: a few microseconds per iteration where five (or three) launches
// per iteration take tens.  Thread t owns rows t, t + 1024, ...; sums go
// through a fixed wave butterfly and a fixed pass over the 16 wave sums, so
// every thread holds the same bits and takes the same branches.  The final
// scalars go to the CGScal the host reads.
// ---------------------------------------------------------------------------
constexpr int kSmallPcgThreads = 1024;

__device__ __forceinline__ double block_sum_1024(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kSmallPcgThreads / 64; w++) s += sh[w];
  __syncthreads();
  return s;
}

template <int BW>
__device__ __forceinline__ double small_precond(int i, const double* __restrict__ r, const double* __restrict__ binv,
                                                const int* __restrict__ bfirst,
                                                const unsigned char* __restrict__ bsize) {
  if constexpr (BW == 1) {
    return binv[i] * r[i];
  } else {
    const int f = bfirst[i], m = bsize[i];
    double zi = 0.0;
#pragma unroll
    for (int j = 0; j < BW; j++)
      if (j < m) zi += binv[(long long)i * BW + j] * r[f + j];
    return zi;
  }
}

template <int BW>
__global__ __launch_bounds__(kSmallPcgThreads) void k_pcg_small(
    int n, const int* __restrict__ rowptr, const int* __restrict__ colind, const double* __restrict__ vals,
    const double* __restrict__ b, double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
    double* __restrict__ p, double* __restrict__ q, const double* __restrict__ binv, const int* __restrict__ bfirst,
    const unsigned char* __restrict__ bsize, double tol2, int maxit, CGScal* __restrict__ S) {
  __shared__ double sh[kSmallPcgThreads / 64];
  const int t = threadIdx.x;
  constexpr int NT = kSmallPcgThreads;
  for (int i = t; i < n; i += NT) {
    x[i] = 0.0;
    r[i] = b[i];
  }
  __syncthreads();
  double l1 = 0.0, l2 = 0.0;
  for (int i = t; i < n; i += NT) {
    const double zi = small_precond<BW>(i, r, binv, bfirst, bsize);
    z[i] = zi;
    p[i] = zi;
    l1 += r[i] * zi;
    l2 += r[i] * r[i];
  }
  double rz = block_sum_1024(l1, sh);
  const double bb = block_sum_1024(l2, sh);
  double rr = bb, pq = 0.0, alpha = 0.0, beta = 0.0;
  int it = 0, brk = 0;
  bool done = bb == 0.0;
  while (!done) {
    double lpq = 0.0;
    for (int i = t; i < n; i += NT) {
      // sixteen entries' loads in flight at a time, summed in order
      const int rb = rowptr[i], re = rowptr[i + 1];
      double s = 0.0;
      for (int k0 = rb; k0 < re; k0 += 16) {
        int c[16];
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const int k = min(k0 + u, re - 1);
          c[u] = colind[k];
          v[u] = vals[k];
        }
        double pc[16];
#pragma unroll
        for (int u = 0; u < 16; u++) pc[u] = p[c[u]];
#pragma unroll
        for (int u = 0; u < 16; u++)
          if (k0 + u < re) s += v[u] * pc[u];
      }
      q[i] = s;
      lpq += p[i] * s;
    }
    pq = block_sum_1024(lpq, sh);
    if (pq == 0.0 || !isfinite(pq)) {
      brk = 1;
      break;
    }
    alpha = rz / pq;
    for (int i = t; i < n; i += NT) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    __syncthreads();
    l1 = 0.0;
    l2 = 0.0;
    for (int i = t; i < n; i += NT) {
      const double zi = small_precond<BW>(i, r, binv, bfirst, bsize);
      z[i] = zi;
      l1 += r[i] * zi;
      l2 += r[i] * r[i];
    }
    const double rzn = block_sum_1024(l1, sh);
    rr = block_sum_1024(l2, sh);
    it++;
    beta = rzn / rz;
    rz = rzn;
    if (rr <= tol2 * bb || it >= maxit) break;
    for (int i = t; i < n; i += NT) p[i] = z[i] + beta * p[i];
    __syncthreads();
  }
  if (t == 0) {
    S->rzp[0] = S->rzp[1] = rz;
    S->pq = pq;
    S->rr = rr;
    S->bb = bb;
    S->alpha = alpha;
    S->beta = beta;
    S->tol2 = tol2;
    S->maxit = maxit;
    S->iterp[0] = S->iterp[1] = it;
    S->fresh = 0;
    S->breakdown = brk;
    S->converged = 1;
  }
}

// ---------------------------------------------------------------------------
// Band-Cholesky preconditioner of small, narrow-banded systems.  In the band
// order k = pinv[e] of the equations (the numbering itself, or a coordinate
// ordering when that is narrower) the matrix has half bandwidth bw <= 63, so
// one 64-lane wavefront holds a whole pivot window.  Layouts (ld = bw + 1):
//   A  the matrix: row k holds A(k, k-bw .. k) at offsets 0 .. bw (the
//      diagonal at bw);
//   L  the factor by columns: row j holds U(j + i, j) at offset i;
//   Lr the factor by rows: row k holds U(k, k - i) at offset i;
// where A = U D^2 U^T (U = L D^-1 unit lower, D = diag L); in both, offset
// 0 holds 1 / L(j, j)^2.  The sweeps then divide nowhere on their chain.
// The factorization runs once per assembled matrix; every application of
// the preconditioner is a forward and a backward sweep by wave 0 of the
// solving workgroup, fed through LDS by the other waves.
// ---------------------------------------------------------------------------
// A workgroup barrier that orders LDS only: unlike __syncthreads it does not
// wait for this wave's outstanding global stores, which no other wave reads.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ double lane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(256) void k_band_fill(long long n, double* __restrict__ a) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_band_gather(int n, int bw, const int* __restrict__ rowptr,
                                                     const int* __restrict__ colind, const double* __restrict__ vals,
                                                     const int* __restrict__ pinv, double* __restrict__ A) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int rk = pinv[r], ld = bw + 1;
  for (int q = rowptr[r]; q < rowptr[r + 1]; q++) {
    const int ck = pinv[colind[q]];
    if (ck <= rk && rk - ck <= bw) A[(long long)rk * ld + (ck - rk + bw)] = vals[q];
  }
}

// Right-looking band Cholesky by one wavefront: at pivot j lane i holds
// row j + i of the reduced window in w[0 .. BW] (w[c] = A'(j + i, j + c)).
// A non-positive pivot sets *bad (the matrix is not positive definite).
template <int BW>
__global__ __launch_bounds__(64) void k_band_factor(int n, int bw, const double* __restrict__ A,
                                                    double* __restrict__ L, double* __restrict__ Lr,
                                                    int* __restrict__ bad) {
  const int lane = threadIdx.x;
  const int ld = bw + 1;
  const bool act = lane <= bw;
  double w[BW + 1];
#pragma unroll
  for (int c = 0; c <= BW; c++) w[c] = 0.0;
  if (act && lane < n) {
    const double* row = A + (long long)lane * ld + (bw - lane);
#pragma unroll
    for (int c = 0; c <= BW; c++)
      if (c <= lane) w[c] = row[c];
  }
  int fail = 0;
  for (int j = 0; j < n; j++) {
    // the row entering the window after this pivot, loaded ahead
    const int r = j + 1 + bw;
    const double nrow = (act && r < n) ? A[(long long)r * ld + lane] : 0.0;
    const double a00 = lane_d(w[0], 0);
    if (!(a00 > 0.0)) fail = 1;
    const double d = sqrt(a00 > 0.0 ? a00 : 1.0);
    const double rd = 1.0 / d;
    const double l = lane == 0 ? d : (act ? w[0] * rd : 0.0);
    if (act && j + lane < n) {
      const double o = lane == 0 ? rd * rd : l * rd;
      L[(long long)j * ld + lane] = o;
      Lr[(long long)(j + lane) * ld + lane] = o;
    }
#pragma unroll
    for (int c = 1; c <= BW; c++) w[c] -= l * lane_d(l, c);
#pragma unroll
    for (int c = 0; c < BW; c++) w[c] = __shfl_down(w[c + 1], 1, 64);
    w[BW] = 0.0;
#pragma unroll
    for (int c = 0; c <= BW; c++) {
      const double v = lane_d(nrow, c);
      if (lane == bw) w[c] = v;
    }
  }
  if (lane == 0 && fail) bad[0] = 1;
}

// 1 / sqrt(a) for a > 0: the hardware estimate and two Newton steps
// (y <- y (1.5 - a y^2 / 2)), accurate to rounding.
__device__ __forceinline__ double rsqrt_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
  const double h = 0.5 * a;
  y = y * fma(-h, y * y, 1.5);
  y = y * fma(-h, y * y, 1.5);
  return y;
}

// Band Cholesky over four wavefronts, for bw <= 47.  Lane l holds the window
// row congruent to l mod 64, so no value moves between lanes; each row's
// entries sit in registers indexed by column, in a frame of 128 columns
// (j0 - 48 .. j0 + 79 for the chunk of 16 pivots starting at j0) that moves
// 16 columns on after each chunk, so every register index is a
// compile-time constant.  The frame is dealt out in panels of 4 columns:
// wave v holds frame columns rho with (rho / 4) mod 4 = v, does every update
// of them in pivot order, and factors its panels itself -- four pivots in a
// row within one wavefront, the multipliers broadcast by readlane.  It then
// puts the panel's four multiplier vectors in LDS (rotated so that distance
// m from each pivot row is element m); after one barrier every wave applies
// the rank-4 update to its own columns, the owner of the next panel first
// to that panel's columns, which it then factors.  Entries above a row's
// diagonal receive updates that are never read.  After each chunk the 16
// finished lanes take the rows 64 further on, staged in LDS two chunks
// ahead (each thread loads its share at the start of a chunk and stores it
// to LDS at the end).  Outputs as k_band_factor: U by columns (L) and by
// rows (Lr), 1 / L(j, j)^2 at offset 0.
template <int BW, int WV>
__device__ __forceinline__ void band_factor_wave(int n, int bw, const double* __restrict__ A,
                                                 double* __restrict__ L, double* __restrict__ Lr,
                                                 int* __restrict__ bad, double (*__restrict__ sA)[16 * 64],
                                                 double (*__restrict__ shl)[4][64], int nch) {
  constexpr int NS = 32, OFF = 48;
  static_assert(BW <= 44, "the column frame holds half bandwidths up to 44");
  // frame column rho lives in w[loc(rho)] of wave own(rho)
  auto own = [](int rho) { return (rho >> 2) & 3; };
  auto loc = [](int rho) { return ((rho >> 4) << 2) | (rho & 3); };
  const int lane = threadIdx.x & 63;
  const int ld = bw + 1;
  double w[NS];
  int row = lane, fail = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int rho = 16 * (s >> 2) + 4 * WV + (s & 3);
    const int col = rho - OFF, e = col - lane + bw;
    w[s] = (col >= 0 && e >= 0 && e <= bw && lane < n) ? A[(long long)lane * ld + e] : 0.0;
  }
  // factor the panel of pivots jp .. jp + 3 at frame columns rho0 .. rho0 + 3
  // (this wave's), its multipliers into buf
  auto panel = [&](int jp, int rho0, double (*buf)[64]) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int j = jp + b;
      if (j < n) {
        const int p = j & 63;
        const double a = w[loc(rho0 + b)];
        const double a00 = lane_d(a, p);
        if (!(a00 > 0.0)) fail = 1;
        const double rd = rsqrt_nr(a00 > 0.0 ? a00 : 1.0);
        const int i = row - j;
        const bool inwin = (unsigned)i <= (unsigned)bw && row < n;
        const double l = inwin ? a * rd : 0.0;
        buf[b][(lane - p) & 63] = l;
#pragma unroll
        for (int d = 1; b + d < 4; d++) w[loc(rho0 + b + d)] -= l * lane_d(l, (p + d) & 63);
        if (inwin) {
          const double o = i == 0 ? rd * rd : l * rd;
          L[(long long)j * ld + i] = o;
          Lr[(long long)row * ld + i] = o;
        }
      } else {
        buf[b][lane] = 0.0;
      }
    }
  };
  // rank-4 update of this wave's columns in [lo, hi] by the panel at frame
  // columns rho0 .. rho0 + 3 (multipliers B, the lanes' own lr)
  auto apply = [&](int rho0, int lo, int hi, const double (*B)[64], const double* lr) {
#pragma unroll
    for (int rho = lo; rho <= hi; rho++) {
      if (own(rho) == WV) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int m = rho - (rho0 + b);
          if (m >= 1 && m <= BW + 3) w[loc(rho)] -= lr[b] * B[b][m];
        }
      }
    }
  };
  // the rows chunk c's finished lanes take, 16c + 64 .. 16c + 79: this
  // thread's four of the 16 x 64 staged values
  const int tid = threadIdx.x;
  double pre[4];
  auto fetch = [&](int c) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = tid + 256 * k, t = e >> 6, q = e & 63, R = 16 * c + 64 + t;
      pre[k] = (c < nch && R < n && q <= bw) ? A[(long long)R * ld + q] : 0.0;
    }
  };
  auto store = [&](int c) {
#pragma unroll
    for (int k = 0; k < 4; k++) sA[c % 3][tid + 256 * k] = pre[k];
  };
  fetch(0);
  store(0);
  fetch(1);
  store(1);
  if (WV == 0) panel(0, OFF, shl[0]);
  lds_barrier();
  for (int c = 0; c < nch; c++) {
    const int j0 = 16 * c;
    fetch(c + 2);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int jp = j0 + 4 * q, rho0 = OFF + 4 * q;
      const double(*B)[64] = shl[q & 1];
      double lr[4];
#pragma unroll
      for (int b = 0; b < 4; b++) lr[b] = B[b][(lane - ((jp + b) & 63)) & 63];
      const int hi = rho0 + 3 + BW;
      if (q < 3) {
        if (own(rho0 + 4) == WV) {
          apply(rho0, rho0 + 4, rho0 + 7, B, lr);
          if (jp + 4 < n) panel(jp + 4, rho0 + 4, shl[(q + 1) & 1]);
          apply(rho0, rho0 + 8, hi, B, lr);
        } else {
          apply(rho0, rho0 + 4, hi, B, lr);
        }
      } else {
        apply(rho0, rho0 + 4, hi, B, lr);
        store(c + 2);
        // the lanes finished in this chunk take their next rows; the frame
        // moves 16 columns (4 of this wave's slots) on
        const bool sw = row >= j0 && row < j0 + 16;
        if (sw) row += 64;
        const int t = row - (j0 + 64);
        const double* S = sA[c % 3];
#pragma unroll
        for (int s = 0; s < NS; s++) {
          const int rho = 16 * (s >> 2) + 4 * WV + (s & 3);
          if (rho >= 112 - BW) {
            const int e = rho - 112 - t + bw;
            const bool ok = sw && e >= 0 && e <= bw;
            const double v = S[ok ? t * 64 + e : 0];
            if (ok) w[s] = v;
          }
        }
#pragma unroll
        for (int s = 0; s < NS - 4; s++) w[s] = w[s + 4];
#pragma unroll
        for (int s = NS - 4; s < NS; s++) w[s] = 0.0;
        if (WV == 0 && j0 + 16 < n) panel(j0 + 16, OFF, shl[0]);
      }
      lds_barrier();
    }
  }
  if (lane == 0 && fail) bad[0] = 1;
}

template <int BW>
__global__ __launch_bounds__(256) void k_band_factor_w(int n, int bw, const double* __restrict__ A,
                                                      double* __restrict__ L, double* __restrict__ Lr,
                                                      int* __restrict__ bad) {
  __shared__ double sA[3][16 * 64];
  __shared__ double shl[2][4][64];
  const int nch = (n + 15) >> 4;
  switch (threadIdx.x >> 6) {
    case 0: band_factor_wave<BW, 0>(n, bw, A, L, Lr, bad, sA, shl, nch); break;
    case 1: band_factor_wave<BW, 1>(n, bw, A, L, Lr, bad, sA, shl, nch); break;
    case 2: band_factor_wave<BW, 2>(n, bw, A, L, Lr, bad, sA, shl, nch); break;
    default: band_factor_wave<BW, 3>(n, bw, A, L, Lr, bad, sA, shl, nch); break;
  }
}

// Sweeps of the band factor, (U D^2 U^T)^-1 t, in blocks of 64 rows.  Lane
// l of wave 0 holds the pending value of the window row congruent to l mod
// 64, so no value moves between lanes: at pivot j the pending value of lane
// j mod 64 is final (U has a unit diagonal), it is broadcast by readlane and
// every other lane subtracts its entry of U times it -- one readlane and one
// FMA on the chain per row.  The finished lanes of each group of CH steps
// (16 when bw <= 48: a finished lane's entries stay zero for the rest of its
// group) write their value (scaled by D^-2 after the forward sweep) and take
// the row 64 further on.  The entries of each block, the t values the
// lanes take next and, forward, the block's D^-2 are staged in LDS by waves
// 1..15 two blocks ahead, as a 64 x 64 tile already rotated so that at step
// u lane l reads element u * 64 + l (zero at distance 0 and past bw or n).
constexpr int kBandTile = 64 * 64 + 128;  // doubles per staged block

// Waves 1..15 stage block b of the factor M (L for the forward sweep, Lr for
// the backward), the t values of rows tr0 .. tr0 + 63 and, forward, the
// block's D^-2 into S.  Step u, lane l reads the entry at distance
// (l - u) mod 64 (forward) or (u - l) mod 64 (backward) of row 64 b + u.
__device__ __forceinline__ void band_stage(int n, int bw, const double* __restrict__ M,
                                           const double* __restrict__ t, int b, int tr0, bool fwd,
                                           double* __restrict__ S) {
  constexpr int NL = kSmallPcgThreads - 64;
  constexpr int K = (kBandTile + NL - 1) / NL;
  const int ld = bw + 1;
  const int e0 = threadIdx.x - 64;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int e = e0 + k * NL;
    double x = 0.0;
    if (e < 64 * 64) {
      const int u = e >> 6, l = e & 63, row = 64 * b + u;
      const int i = fwd ? ((l - u) & 63) : ((u - l) & 63);
      if (i >= 1 && i <= bw && row < n) x = M[(long long)row * ld + i];
    } else if (e < 64 * 64 + 64) {
      const int row = tr0 + (e - 64 * 64);
      if (row >= 0 && row < n) x = t[row];
    } else if (e < kBandTile) {
      const int row = 64 * b + (e - 64 * 64 - 64);
      if (fwd && row < n) x = M[(long long)row * ld];
    }
    v[k] = x;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int e = e0 + k * NL;
    if (e < kBandTile) S[e] = v[k];
  }
}

// One staged block of a sweep by wave 0: 64 steps (forward: u = 0 .. 63,
// backward: u = 63 .. 0); rw the lanes' pending values; returns the block's
// finished values (lane l: row 64 b + l), scaled by D^-2 when forward.
template <bool FWD, int CH>
__device__ __forceinline__ double band_block(const double* __restrict__ S, double& rw) {
  const int lane = threadIdx.x & 63;
  const double tn = S[64 * 64 + lane];
  const double d2 = FWD ? S[64 * 64 + 64 + lane] : 1.0;
  double out = 0.0;
#pragma unroll 1
  for (int c = 0; c < 4; c++) {
    const int u0 = FWD ? 16 * c : 63 - 16 * c;
    double lv[16];
#pragma unroll
    for (int k = 0; k < 16; k++) lv[k] = S[(FWD ? u0 + k : u0 - k) * 64 + lane];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int u = FWD ? u0 + k : u0 - k;
      rw = fma(-lv[k], lane_d(rw, u), rw);
      if (CH == 1 || k % CH == CH - 1) {
        // the group's lanes are final: keep them, take the next rows
        const int g0 = FWD ? u - (CH - 1) : u;
        if (lane >= g0 && lane < g0 + CH) {
          out = FWD ? rw * d2 : rw;
          rw = tn;
        }
      }
    }
  }
  return out;
}

// t <- (U D^2 U^T)^-1 t in band order.  Called by all kSmallPcgThreads threads
// of the workgroup (it holds barriers); stage: 3 * kBandTile doubles of LDS.
__device__ __attribute__((noinline)) void band_sweeps(int n, int bw, const double* __restrict__ L,
                                                      const double* __restrict__ Lr, double* __restrict__ t,
                                                      double* __restrict__ stage) {
  const int tid = threadIdx.x, lane = tid & 63;
  const bool w0 = tid < 64;
  const int nb = (n + 63) >> 6;
  // forward: L y = t; block b's staged t values are rows 64 (b + 1) + l
  double rw = 0.0;
  if (w0) {
    rw = lane < n ? t[lane] : 0.0;
  } else {
    band_stage(n, bw, L, t, 0, 64, true, stage);
    if (nb > 1) band_stage(n, bw, L, t, 1, 128, true, stage + kBandTile);
  }
  lds_barrier();
  for (int b = 0; b < nb; b++) {
    if (w0) {
      double* S = stage + (b % 3) * kBandTile;
      const double y = bw <= 48 ? band_block<true, 16>(S, rw) : band_block<true, 1>(S, rw);
      if (64 * b + lane < n) t[64 * b + lane] = y;
    } else if (b + 2 < nb) {
      band_stage(n, bw, L, t, b + 2, 64 * (b + 3), true, stage + ((b + 2) % 3) * kBandTile);
    }
    lds_barrier();
  }
  // the backward staging reads the forward sweep's t from global memory
  __syncthreads();
  // backward: L^T x = y, blocks from the last; block b's staged t values are
  // rows 64 (b - 1) + l
  rw = 0.0;
  if (w0) {
    const int r = 64 * (nb - 1) + lane;
    rw = r < n ? t[r] : 0.0;
  } else {
    band_stage(n, bw, Lr, t, nb - 1, 64 * (nb - 2), false, stage);
    if (nb > 1) band_stage(n, bw, Lr, t, nb - 2, 64 * (nb - 3), false, stage + kBandTile);
  }
  __syncthreads();
  for (int q = 0; q < nb; q++) {
    const int b = nb - 1 - q;
    if (w0) {
      double* S = stage + (q % 3) * kBandTile;
      const double x = bw <= 48 ? band_block<false, 16>(S, rw) : band_block<false, 1>(S, rw);
      if (64 * b + lane < n) t[64 * b + lane] = x;
    } else if (q + 2 < nb) {
      band_stage(n, bw, Lr, t, b - 2, 64 * (b - 3), false, stage + ((q + 2) % 3) * kBandTile);
    }
    lds_barrier();
  }
  __syncthreads();
}

// The whole PCG solve of a band-preconditioned system in one workgroup:
// the product and the vector updates over all threads, the preconditioner
// by wave 0 between barriers (t: a band-ordered work vector).  With the
// exact factor the first iteration meets the tolerance; the test comes
// before the next application, so a converged solve applies it once.
__global__ __launch_bounds__(kSmallPcgThreads) void k_pcg_band(
    int n, const int* __restrict__ rowptr, const int* __restrict__ colind, const double* __restrict__ vals, int bw,
    const int* __restrict__ pinv, const double* __restrict__ L, const double* __restrict__ Lr, double tol2, int maxit,
    BandRhs r0, BandRhs r1) {
  // workgroup k solves right-hand side k (two independent solves with one
  // factor run side by side on two CUs)
  const BandRhs& R = blockIdx.x == 0 ? r0 : r1;
  const double* __restrict__ b = R.b;
  double* __restrict__ x = R.x;
  double* __restrict__ r = R.r;
  double* __restrict__ z = R.z;
  double* __restrict__ p = R.p;
  double* __restrict__ q = R.q;
  double* __restrict__ t = R.t;
  CGScal* __restrict__ S = R.S;
  __shared__ double sh[kSmallPcgThreads / 64];
  __shared__ double stage[3 * kBandTile];
  const int tid = threadIdx.x;
  constexpr int NT = kSmallPcgThreads;
  auto precond = [&]() {
    for (int i = tid; i < n; i += NT) t[pinv[i]] = r[i];
    __syncthreads();
    band_sweeps(n, bw, L, Lr, t, stage);
    for (int i = tid; i < n; i += NT) z[i] = t[pinv[i]];
    __syncthreads();
  };
  double l1 = 0.0, l2 = 0.0, rz = 0.0, bb = 0.0, rr = 0.0;
  bool done;
  if (!R.warm) {
    for (int i = tid; i < n; i += NT) {
      x[i] = 0.0;
      r[i] = b[i];
    }
    __syncthreads();
    precond();
    for (int i = tid; i < n; i += NT) {
      p[i] = z[i];
      l1 += r[i] * z[i];
      l2 += r[i] * r[i];
    }
    rz = block_sum_1024(l1, sh);
    bb = block_sum_1024(l2, sh);
    rr = bb;
    done = bb == 0.0;
  } else {
    // from the guess in x: r = b - A x; a guess already within the
    // tolerance ends the solve before the first iteration
    double lb = 0.0;
    for (int i = tid; i < n; i += NT) {
      double s = 0.0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) s += vals[k] * x[colind[k]];
      const double ri = b[i] - s;
      r[i] = ri;
      l2 += ri * ri;
      lb += b[i] * b[i];
    }
    bb = block_sum_1024(lb, sh);
    rr = block_sum_1024(l2, sh);
    done = bb == 0.0 || rr <= tol2 * bb;
    if (!done) {
      precond();
      for (int i = tid; i < n; i += NT) {
        p[i] = z[i];
        l1 += r[i] * z[i];
      }
      rz = block_sum_1024(l1, sh);
    }
  }
  double pq = 0.0, alpha = 0.0, beta = 0.0;
  int it = 0, brk = 0;
  while (!done) {
    double lpq = 0.0;
    for (int i = tid; i < n; i += NT) {
      const int rb = rowptr[i], re = rowptr[i + 1];
      double s = 0.0;
      for (int k0 = rb; k0 < re; k0 += 16) {
        int c[16];
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const int k = min(k0 + u, re - 1);
          c[u] = colind[k];
          v[u] = vals[k];
        }
        double pc[16];
#pragma unroll
        for (int u = 0; u < 16; u++) pc[u] = p[c[u]];
#pragma unroll
        for (int u = 0; u < 16; u++)
          if (k0 + u < re) s += v[u] * pc[u];
      }
      q[i] = s;
      lpq += p[i] * s;
    }
    pq = block_sum_1024(lpq, sh);
    if (pq == 0.0 || !isfinite(pq)) {
      brk = 1;
      break;
    }
    alpha = rz / pq;
    l2 = 0.0;
    for (int i = tid; i < n; i += NT) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      r[i] = ri;
      l2 += ri * ri;
    }
    rr = block_sum_1024(l2, sh);
    it++;
    if (rr <= tol2 * bb || it >= maxit) break;
    precond();
    l1 = 0.0;
    for (int i = tid; i < n; i += NT) l1 += r[i] * z[i];
    const double rzn = block_sum_1024(l1, sh);
    beta = rzn / rz;
    rz = rzn;
    for (int i = tid; i < n; i += NT) p[i] = z[i] + beta * p[i];
    __syncthreads();
  }
  if (tid == 0) {
    S->rzp[0] = S->rzp[1] = rz;
    S->pq = pq;
    S->rr = rr;
    S->bb = bb;
    S->alpha = alpha;
    S->beta = beta;
    S->tol2 = tol2;
    S->maxit = maxit;
    S->iterp[0] = S->iterp[1] = it;
    S->fresh = 0;
    S->breakdown = brk;
    S->converged = 1;
  }
}

// Jacobi / nodal block-Jacobi inverse from the assembled CSR matrix.
// One thread per node block: gathers the m x m diagonal block (m <= M, the
// nodal DOF count), inverts it by Gauss-Jordan with partial pivoting and
// writes its rows.  The block is padded to M x M with identity rows so that
// every index is a compile-time constant (the matrix stays in registers);
// the padding never mixes with the real rows, whose arithmetic is that of
// the unpadded m x m elimination.
template <int M>
__global__ __launch_bounds__(128) void k_block_inverse(int nblk, const int* __restrict__ blk_first,
                                const unsigned char* __restrict__ blk_size,
                                const int* __restrict__ rowptr, const int* __restrict__ colind,
                                const double* __restrict__ vals, double* __restrict__ binv,
                                int* __restrict__ bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const int f = blk_first[k], m = blk_size[k];
  if (M == 1) {
    // Jacobi: each equation its own block
    for (int j = 0; j < m; j++) {
      const int row = f + j;
      double d = 0.0;
      for (int q = rowptr[row]; q < rowptr[row + 1]; q++)
        if (colind[q] == row) d = vals[q];
      if (d == 0.0) bad[0] = 1;
      binv[row] = (d != 0.0) ? 1.0 / d : 0.0;
    }
    return;
  }
  double a[M][2 * M];
#pragma unroll
  for (int i = 0; i < M; i++) {
#pragma unroll
    for (int j = 0; j < 2 * M; j++) a[i][j] = (j - M == i || (i >= m && j == i)) ? 1.0 : 0.0;
    if (i < m) {
      // the block's columns f .. f+m-1 follow each other in the sorted row
      const int row = f + i, e = rowptr[row + 1];
      int q = rowptr[row];
      while (q < e && colind[q] < f) q++;
#pragma unroll
      for (int j = 0; j < M; j++)
        if (j < m && q < e && colind[q] == f + j) a[i][j] = vals[q++];
    }
  }
#pragma unroll
  for (int c = 0; c < M; c++) {
    int piv = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int i = c + 1; i < M; i++) {
      const double v = fabs(a[i][c]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    if (best == 0.0) {
      bad[0] = 1;
      return;
    }
#pragma unroll
    for (int i = c + 1; i < M; i++) {
      const bool sw = (piv == i);
#pragma unroll
      for (int j = 0; j < 2 * M; j++) {
        const double tc = a[c][j], ti = a[i][j];
        a[c][j] = sw ? ti : tc;
        a[i][j] = sw ? tc : ti;
      }
    }
    const double inv = 1.0 / a[c][c];
#pragma unroll
    for (int j = 0; j < 2 * M; j++) a[c][j] *= inv;
#pragma unroll
    for (int i = 0; i < M; i++)
      if (i != c) {
        const double fct = a[i][c];
#pragma unroll
        for (int j = 0; j < 2 * M; j++) a[i][j] -= fct * a[c][j];
      }
  }
#pragma unroll
  for (int i = 0; i < M; i++)
    if (i < m)
#pragma unroll
      for (int j = 0; j < M; j++) binv[(long long)(f + i) * M + j] = (j < m) ? a[i][M + j] : 0.0;
}

// ---------------------------------------------------------------------------
// Small vector kernels
// ---------------------------------------------------------------------------
// U[node dof] += dU[eq] over the DOF map (AnalysisModel::incrDisp)
__global__ void k_incr_disp(long long ndofs, const int* __restrict__ eqmap, const double* __restrict__ dU,
                            double scale, double* __restrict__ U) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndofs) return;
  const int a = eqmap[i];
  if (a >= 0) U[i] += scale * dU[a];
}

// y = a*x + b*y
__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = a * x[i] + b * y[i];
}

// plain copy (HBM traffic calibration of the profiler counters)
__global__ void k_copy(long long n, const double* __restrict__ x, double* __restrict__ y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i];
}

// partial sums of x.y per block
__global__ __launch_bounds__(256) void k_dot_partial(int n, const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ part) {
  __shared__ double sh[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = (i < n) ? x[i] * y[i] : 0.0;
  const double s = block_sum<256>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_sum_partials(int nb, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  __shared__ double sh[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) v += part[i];
  const double s = block_sum<256>(v, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_assemble_csr(hipStream_t s, long long nnz, const int* ptr, const int* idx, const double* Ke,
                         double* vals) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(k_assemble_csr, dim3(ceil_div(nnz, 256)), dim3(256), 0, s, nnz, ptr, idx, Ke, vals);
}
void launch_assemble_rhs(hipStream_t s, int neq, const int* ptr, const int* idx, const double* Re,
                         double lambda, const double* P, double* B) {
  if (neq <= 0) return;
  hipLaunchKernelGGL(k_assemble_rhs, dim3(ceil_div(neq, 256)), dim3(256), 0, s, neq, ptr, idx, Re, lambda,
                     P, B);
}
void launch_assemble_rhs_dyn(hipStream_t s, int neq, const int* ptr, const int* idx, const double* Re,
                             double lambda, const double* P, const double* Meq, const double* Req, double ag,
                             double alphaM, const int* eq2dof, const double* V, const double* A, double* B) {
  if (neq <= 0) return;
  hipLaunchKernelGGL(k_assemble_rhs_dyn, dim3(ceil_div(neq, 256)), dim3(256), 0, s, neq, ptr, idx, Re, lambda, P,
                     Meq, Req, ag, alphaM, eq2dof, V, A, B);
}
void launch_newmark_predict(hipStream_t s, long long ndofs, const int* eqmap, const double* Uc, const double* Vc,
                            const double* Ac, double a1, double a2, double a3, double a4, double* U, double* V,
                            double* A) {
  if (ndofs <= 0) return;
  hipLaunchKernelGGL(k_newmark_predict, dim3(ceil_div(ndofs, 256)), dim3(256), 0, s, ndofs, eqmap, Uc, Vc, Ac, a1,
                     a2, a3, a4, U, V, A);
}
void launch_newmark_correct(hipStream_t s, long long ndofs, const int* eqmap, const double* dU, double c2, double c3,
                            double* U, double* V, double* A) {
  if (ndofs <= 0) return;
  hipLaunchKernelGGL(k_newmark_correct, dim3(ceil_div(ndofs, 256)), dim3(256), 0, s, ndofs, eqmap, dU, c2, c3, U, V,
                     A);
}
void launch_add_diag(hipStream_t s, int n, const long long* pos, const double* m, double coef,
                     double* vals) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_diag, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, pos, m, coef, vals);
}
// grid of the CG kernels: enough workgroups to fill the chip, few enough
// that the single-workgroup reductions of their partials stay short.  The
// SpMV, whose rows vary in length, balances better over more workgroups.
constexpr int kCgMaxBlocks = 2048;
constexpr int kSpmvMaxBlocks = 8192;
// The symmetric product runs one unit of block rows per workgroup up to this
// many units, so that the workgroups resident on an XCD sweep one front of
// its rows (beyond it a workgroup takes a contiguous run of units).
constexpr int kSymMaxBlocks = 65536;
int spmv_blocks(int n, int lanes) { return std::min(ceil_div(n, 256 / lanes), kSpmvMaxBlocks); }
static int vec_blocks(int n) { return std::min(ceil_div(n, 256), kCgMaxBlocks); }
// the fused alpha / beta kernels: at most this many partials per reduction;
// the update's partials then start at part + kCgFuseMax
constexpr int kCgFuseMax = 1024;
size_t cg_part_size(int n) {
  n = std::max(n, 1);
  // the symmetric product's workgroups: at most one per 8 equations (block
  // rows of >= 2 equations, >= 4 block rows per workgroup)
  const int sym = std::min(ceil_div(n, 8), kSymMaxBlocks);
  return 3 * (size_t)std::max({ceil_div(n, 256), spmv_blocks(n, 64), sym, kCgFuseMax}) + kCgFuseMax + 16;
}
// SpMV launches; otf != nullptr: the fused PCG's product with the new
// direction formed on the fly (OtfArgs)
template <bool OTF, int M>
static void launch_spmv_nb_t(hipStream_t s, int lanes, int n, const int* rowptr, const int* nbcol,
                             const double* vals, const double* x, double* y, double* partial, const int* stop,
                             const OtfArgs& o) {
  const dim3 g(spmv_blocks(n, lanes)), b(256);
  switch (lanes) {
#define XG_NB(L)                                                                                           \
  case L:                                                                                                 \
    hipLaunchKernelGGL((k_spmv_nb_dot<L, M, OTF>), g, b, 0, s, n, rowptr, nbcol, vals, x, y, partial, stop, o); \
    break;
    XG_NB(4) XG_NB(8) XG_NB(16) XG_NB(32)
#undef XG_NB
    default: hipLaunchKernelGGL((k_spmv_nb_dot<64, M, OTF>), g, b, 0, s, n, rowptr, nbcol, vals, x, y, partial, stop,
                                o); break;
  }
}

template <bool OTF, int M>
static void launch_spmv_sym_t(hipStream_t s, const SymSpmv& S, const double* x, double* y, double* partial,
                              const int* stop, const OtfArgs& o) {
  const dim3 g(spmv_grid(S.nbr * M, S.lanes, &S)), b(256);
  switch (S.lanes) {
#define XG_SYM(G)                                                                                              \
  case G:                                                                                                     \
    hipLaunchKernelGGL((k_spmv_sym_dot<M, G, OTF>), g, b, 0, s, S.nbr, S.ptr, S.col, S.blk, S.val, x, y, partial, \
                       stop, o);                                                                              \
    break;
    XG_SYM(4) XG_SYM(8) XG_SYM(16) XG_SYM(32) XG_SYM(64)
#undef XG_SYM
    default: throw Error(XG_ERR_ARG, "symmetric SpMV lanes must be 4, 8, 16, 32 or 64");
  }
}

template <bool OTF>
static void launch_spmv_t(hipStream_t s, int lanes, int n, const int* rowptr, const int* colind, const double* vals,
                          const double* x, double* y, double* partial, const int* stop, int nbm, const int* nbcol,
                          const SymSpmv* sym, const OtfArgs& o) {
  if (n <= 0) return;
  if (sym && sym->m) {
    if (sym->m == 2) return launch_spmv_sym_t<OTF, 2>(s, *sym, x, y, partial, stop, o);
    if (sym->m == 3) return launch_spmv_sym_t<OTF, 3>(s, *sym, x, y, partial, stop, o);
    throw Error(XG_ERR_ARG, "symmetric SpMV block size must be 2 or 3");
  }
  switch (nbm) {
    case 2: return launch_spmv_nb_t<OTF, 2>(s, lanes, n, rowptr, nbcol, vals, x, y, partial, stop, o);
    case 3: return launch_spmv_nb_t<OTF, 3>(s, lanes, n, rowptr, nbcol, vals, x, y, partial, stop, o);
    case 6: return launch_spmv_nb_t<OTF, 6>(s, lanes, n, rowptr, nbcol, vals, x, y, partial, stop, o);
    default: break;
  }
  const dim3 g(spmv_blocks(n, lanes)), b(256);
  switch (lanes) {
#define XG_CSR(L)                                                                                       \
  case L:                                                                                              \
    hipLaunchKernelGGL((k_spmv_dot<L, OTF>), g, b, 0, s, n, rowptr, colind, vals, x, y, partial, stop, o); \
    break;
    XG_CSR(4) XG_CSR(8) XG_CSR(16) XG_CSR(32)
#undef XG_CSR
    default: hipLaunchKernelGGL((k_spmv_dot<64, OTF>), g, b, 0, s, n, rowptr, colind, vals, x, y, partial, stop, o);
             break;
  }
}

int spmv_grid(int n, int lanes, const SymSpmv* sym) {
  if (sym && sym->m) return std::min(ceil_div(std::max(sym->nbr, 1), 256 / lanes), kSymMaxBlocks);
  return spmv_blocks(n, lanes);
}

void launch_sym_convert(hipStream_t s, const SymSpmv& S, int nup, const int* src, const int* len,
                        const double* vals, double* sval) {
  if (nup <= 0) return;
  const dim3 g(std::min(ceil_div(nup, 256), 65536)), b(256);
  if (S.m == 2) hipLaunchKernelGGL(k_sym_convert<2>, g, b, 0, s, nup, src, len, vals, sval);
  else if (S.m == 3) hipLaunchKernelGGL(k_sym_convert<3>, g, b, 0, s, nup, src, len, vals, sval);
  else throw Error(XG_ERR_ARG, "symmetric SpMV block size must be 2 or 3");
}

__global__ __launch_bounds__(256) void k_vals_differ(long long n, const double* __restrict__ a,
                                                     const double* __restrict__ b, int* __restrict__ flag) {
  bool d = false;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    d |= __double_as_longlong(a[i]) != __double_as_longlong(b[i]);
  if (d) *flag = 1;  // every writer stores the same value
}
void launch_vals_differ(hipStream_t s, long long n, const double* a, const double* b, int* flag) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_vals_differ, dim3((unsigned)std::min<long long>((n + 255) / 256, 2048)), dim3(256), 0, s, n, a,
                     b, flag);
}

void launch_sweep_perm(hipStream_t s, int nbr, int B, const int* ord, const double* src, double* dst, bool back) {
  if (nbr <= 0 || B <= 0) return;
  const long long total = (long long)nbr * B;
  const dim3 g((unsigned)std::min<long long>((total + 255) / 256, 8192)), b(256);
  hipLaunchKernelGGL(k_sweep_perm, g, b, 0, s, nbr, B, ord, src, dst, back ? 1 : 0);
}

void launch_spmv(hipStream_t s, int lanes, int n, const int* rowptr, const int* colind, const double* vals,
                 const double* x, double* y, double* partial, const int* stop, int nbm, const int* nbcol,
                 const SymSpmv* sym) {
  launch_spmv_t<false>(s, lanes, n, rowptr, colind, vals, x, y, partial, stop, nbm, nbcol, sym, OtfArgs{});
}
void launch_block_inverse(hipStream_t s, int nblk, const int* first, const unsigned char* size,
                          const int* rowptr, const int* colind, const double* vals, int bw, double* binv,
                          int* bad) {
  if (nblk <= 0) return;
  const dim3 g(ceil_div(nblk, 128)), b(128);
  switch (bw) {
    case 1: hipLaunchKernelGGL(k_block_inverse<1>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    case 2: hipLaunchKernelGGL(k_block_inverse<2>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    case 3: hipLaunchKernelGGL(k_block_inverse<3>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    case 4: hipLaunchKernelGGL(k_block_inverse<4>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    case 5: hipLaunchKernelGGL(k_block_inverse<5>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    case 6: hipLaunchKernelGGL(k_block_inverse<6>, g, b, 0, s, nblk, first, size, rowptr, colind, vals, binv, bad); break;
    default: throw Error(XG_ERR_ARG, "block-Jacobi block size must be 1..6");
  }
}
void launch_incr_disp(hipStream_t s, long long ndofs, const int* eqmap, const double* dU, double scale,
                      double* U) {
  if (ndofs <= 0) return;
  hipLaunchKernelGGL(k_incr_disp, dim3(ceil_div(ndofs, 256)), dim3(256), 0, s, ndofs, eqmap, dU, scale, U);
}
void launch_copy(hipStream_t s, long long n, const double* x, double* y) {
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, n, x, y);
}
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, a, x, b, y);
}
void launch_dot(hipStream_t s, int n, const double* x, const double* y, double* part, double* out) {
  const int nb = ceil_div(n > 0 ? n : 1, 256);
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, s, n, x, y, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, part, out);
}

// PCG driver pieces ----------------------------------------------------------
size_t cg_scal_bytes() { return sizeof(CGScal); }

void cg_start(hipStream_t s, int n, const double* b, double* x, double* r, double* z, double* p, int bw,
              const double* binv, const int* bfirst, const unsigned char* bsize, double* part, double tol2,
              int maxit, void* scal, const double* q) {
  const int nb = ceil_div(n, 256);
  hipLaunchKernelGGL(k_cg_start, dim3(nb), dim3(256), 0, s, n, b, q, x, r, z, p, bw, binv, bfirst, bsize, part);
  hipLaunchKernelGGL(k_cg_init_scal, dim3(1), dim3(256), 0, s, nb, part, tol2, maxit, (CGScal*)scal);
}

void cg_iterations_check(int bw, const SymSpmv* sym) {
  if (bw < 1 || bw > 6) throw Error(XG_ERR_ARG, "PCG block width must be 1..6");
  if (sym && sym->m) {
    if (sym->m != 2 && sym->m != 3) throw Error(XG_ERR_ARG, "symmetric SpMV block size must be 2 or 3");
    const int g = sym->lanes;
    if (g != 4 && g != 8 && g != 16 && g != 32 && g != 64)
      throw Error(XG_ERR_ARG, "symmetric SpMV lanes must be 4, 8, 16, 32 or 64");
  }
}

void cg_iterations(hipStream_t s, int iters, int it0, int lanes, int n, const int* rowptr,
                   const int* colind, const double* vals, double* x, double* r0, double* r1, double* z,
                   double* p, double* p2, double* q, int bw, const double* binv, const int* bfirst,
                   const unsigned char* bsize, double* part, void* scal, int nbm, const int* nbcol, bool fuse,
                   const SymSpmv* sym) {
  CGScal* S = (CGScal*)scal;
  const int lanes_mv = sym && sym->m ? sym->lanes : lanes;
  const int nbs = spmv_grid(n, lanes_mv, sym);
  const int nbv = vec_blocks(n);
  const int* stop = &S->converged;
  fuse = fuse && nbs <= kCgFuseMax && nbv <= kCgFuseMax;
  double* part2 = part + kCgFuseMax;
  cg_iterations_check(bw, sym);
  for (int it = 0; it < iters; it++) {
    const int par = (it0 + it) & 1;
    double* ri = par ? r1 : r0;
    double* ro = par ? r0 : r1;
    if (fuse) {
      // two launches: the product forming the new direction on the fly
      // (p alternates between p and p2), then the update with alpha
      double* pold = par ? p2 : p;
      double* pnew = par ? p : p2;
      const OtfArgs o{z, pold, pnew, S, par, nbv, part2};
      launch_spmv_t<true>(s, lanes, n, rowptr, colind, vals, pold, q, part, stop, nbm, nbcol, sym, o);
      switch (bw) {
#define XG_CGU(W)                                                                                                  \
  case W:                                                                                                         \
    hipLaunchKernelGGL(k_cg_update_fused<W>, dim3(nbv), dim3(256), 0, s, n, S, par, nbs, part, pnew, q, x, ri, ro, z, \
                       binv, bfirst, bsize, part2);                                                               \
    break;
        XG_CGU(1) XG_CGU(2) XG_CGU(3) XG_CGU(4) XG_CGU(5) XG_CGU(6)
#undef XG_CGU
      }
    } else {
      launch_spmv(s, lanes, n, rowptr, colind, vals, p, q, part, stop, nbm, nbcol, sym);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, nbs, part, S, par);
      switch (bw) {
#define XG_CGU(W)                                                                                              \
  case W:                                                                                                     \
    hipLaunchKernelGGL(k_cg_update<W>, dim3(nbv), dim3(256), 0, s, n, S, p, q, x, ri, ro, z, binv, bfirst, bsize, \
                       part);                                                                                 \
    break;
        XG_CGU(1) XG_CGU(2) XG_CGU(3) XG_CGU(4) XG_CGU(5) XG_CGU(6)
#undef XG_CGU
      }
      hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, s, nbv, part, S, par);
      hipLaunchKernelGGL(k_cg_newp, dim3(nbv), dim3(256), 0, s, n, S, z, p);
    }
  }
}

void band_factor(hipStream_t s, int n, int bw, const int* rowptr, const int* colind, const double* vals,
                 const int* pinv, double* A, double* L, double* Lr, int* bad) {
  if (n <= 0) return;
  if (bw < 0 || bw > 63) throw Error(XG_ERR_ARG, "band preconditioner: half bandwidth must be 0..63");
  const long long sz = (long long)n * (bw + 1);
  const int fg = (int)std::min<long long>((sz + 255) / 256, 4096);
  hipLaunchKernelGGL(k_band_fill, dim3(fg), dim3(256), 0, s, sz, A);
  hipLaunchKernelGGL(k_band_fill, dim3(fg), dim3(256), 0, s, sz, L);
  hipLaunchKernelGGL(k_band_fill, dim3(fg), dim3(256), 0, s, sz, Lr);
  hipLaunchKernelGGL(k_band_gather, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, bw, rowptr, colind, vals, pinv, A);
  const dim3 g(1), b(256);
  if (bw <= 15) hipLaunchKernelGGL(k_band_factor_w<15>, g, b, 0, s, n, bw, A, L, Lr, bad);
  else if (bw <= 23) hipLaunchKernelGGL(k_band_factor_w<23>, g, b, 0, s, n, bw, A, L, Lr, bad);
  else if (bw <= 31) hipLaunchKernelGGL(k_band_factor_w<31>, g, b, 0, s, n, bw, A, L, Lr, bad);
  else if (bw <= 39) hipLaunchKernelGGL(k_band_factor_w<39>, g, b, 0, s, n, bw, A, L, Lr, bad);
  else if (bw <= 44) hipLaunchKernelGGL(k_band_factor_w<44>, g, b, 0, s, n, bw, A, L, Lr, bad);
  else hipLaunchKernelGGL(k_band_factor<63>, dim3(1), dim3(64), 0, s, n, bw, A, L, Lr, bad);
}

void cg_solve_band(hipStream_t s, int n, const int* rowptr, const int* colind, const double* vals, int bw,
                   const int* pinv, const double* L, const double* Lr, double tol2, int maxit, const BandRhs& r0,
                   const BandRhs* r1) {
  hipLaunchKernelGGL(k_pcg_band, dim3(r1 ? 2 : 1), dim3(kSmallPcgThreads), 0, s, n, rowptr, colind, vals, bw, pinv, L,
                     Lr, tol2, maxit, r0, r1 ? *r1 : r0);
}

bool cg_small_fits(int n, long long nnz) { return n > 0 && n <= 1024 && nnz <= 32768; }

void cg_solve_small(hipStream_t s, int n, const int* rowptr, const int* colind, const double* vals, const double* b,
                    double* x, double* r, double* z, double* p, double* q, int bw, const double* binv,
                    const int* bfirst, const unsigned char* bsize, double tol2, int maxit, void* scal) {
  CGScal* S = (CGScal*)scal;
  const dim3 g(1), blk(kSmallPcgThreads);
  switch (bw) {
#define XG_PS(W)                                                                                              \
  case W:                                                                                                    \
    hipLaunchKernelGGL(k_pcg_small<W>, g, blk, 0, s, n, rowptr, colind, vals, b, x, r, z, p, q, binv, bfirst, \
                       bsize, tol2, maxit, S);                                                               \
    break;
    XG_PS(1) XG_PS(2) XG_PS(3) XG_PS(4) XG_PS(5) XG_PS(6)
#undef XG_PS
    default: throw Error(XG_ERR_ARG, "PCG block width must be 1..6");
  }
}

void cg_read_scal(hipStream_t s, const void* scal, int* converged, int* iter, double* rr, double* bb,
                  int* breakdown) {
  CGScal h;
  XG_HIP(hipMemcpyAsync(&h, scal, sizeof(CGScal), hipMemcpyDeviceToHost, s));
  XG_HIP(hipStreamSynchronize(s));
  *converged = h.converged;
  *iter = std::max(h.iterp[0], h.iterp[1]);
  *rr = h.rr;
  *bb = h.bb;
  *breakdown = h.breakdown;
}

}  // namespace xg
