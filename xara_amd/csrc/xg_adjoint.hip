// Adjoint kernels of the frame transient on gfx950 (one 64-lane wavefront per
// section, fibers f = lane, lane + 64, ...):
//   k_adj_strain_adjoint -- for the converged step: the strain adjoint of
//       every fiber, J_eps^T hbar (the history adjoint pulled back through
//       the trial update), reduced into a section resultant s_hat such that
//       the element resisting-force kernel turns it into the nodal vector
//       sum_f (d eps_f / dU)^T J_eps^T hbar;
//   k_adj_history -- with the adjoint solution mu: the fiber's virtual
//       strain eps(mu), the weight w = wt L A eps(mu) that mu^T F_int puts on
//       its stress, and the pull-back of hbar - w e_sigma to the committed
//       history (written in place) and to the material parameters
//       (accumulated per fiber);
//   k_adj_rhs / k_adj_newmark -- the Newmark adjoint recursion on the
//       equation vectors; k_loss_* -- the loss and its gradient;
//   k_reduce_mat -- per-material sums of the per-fiber parameter adjoints.
#include "xg_adjoint.h"
#include "xg_frame_impl.h"
#include "xg_material_adj.h"

namespace xg {

template <bool D3>
__global__ __launch_bounds__(256) void k_adj_strain_adjoint(FrameDev f, const double* __restrict__ U,
                                                            const double* __restrict__ hbs,
                                                            const double* __restrict__ hbc,
                                                            double* __restrict__ shat) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long sec = (long long)blockIdx.x * 4 + wave;
  if (sec >= f.nsec) return;
  const int e = (int)(sec / f.nip), ip = (int)(sec % f.nip);
  double d[4];
  section_deformation<D3>(f, sec, U, d);
  const int def = f.secdef[e];
  const int f0 = f.def_off[def], nf = f.def_off[def + 1] - f0;
  const long long sb0 = f.steel_base[sec], cb0 = f.conc_base[sec];
  double r0 = 0, r1 = 0, r2 = 0;
  for (int i = lane; i < nf; i += 64) {
    const int fi = f0 + i;
    const int kind = f.fkind[fi];
    const double y = f.fy[fi];
    const double z = D3 ? f.fz[fi] : 0.0;
    const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
    double eb = 0.0;
    if (kind == 1) {
      const long long n = f.nsteel, idx = sb0 + f.flocal[fi];
      double c[10], tb[10], cbo[10], thb[3];
#pragma unroll
      for (int k = 0; k < 10; k++) c[k] = f.st_c[k * n + idx];
#pragma unroll
      for (int k = 0; k < 9; k++) tb[k] = hbs[k * n + idx];
      tb[9] = 0.0;
      steel02_vjp(f.steelp[f.fmat[fi]], c, f.kon_c[idx], strain, tb, cbo, thb, eb);
    } else if (kind == 2) {
      const long long n = f.nconc, idx = cb0 + f.flocal[fi];
      double c[5], tb[5], cbo[5], fcb;
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = f.cs_c[k * n + idx];
#pragma unroll
      for (int k = 0; k < 4; k++) tb[k] = hbc[k * n + idx];
      tb[4] = 0.0;
      concrete02_vjp(f.concp[f.fmat[fi]], c, strain, tb, cbo, fcb, eb);
    }
    r0 += eb;
    r1 += -y * eb;
    if (D3) r2 += z * eb;
  }
  r0 = wave_sum(r0);
  r1 = wave_sum(r1);
  if (D3) r2 = wave_sum(r2);
  if (lane == 0) {
    // the basic forces are q = L sum_ip wt B_ip^T s_ip (B_ip holding the
    // 1/L of the section deformation), so s_hat = r / (wt L)
    const long long n = f.nsec;
    const double wL = gl5_wt(ip) * 0.5 * f.geo[(long long)e * 10];
    shat[0 * n + sec] = r0 / wL;
    shat[1 * n + sec] = r1 / wL;
    if (D3) {
      shat[2 * n + sec] = r2 / wL;
      shat[3 * n + sec] = 0.0;
    }
  }
}

template <bool D3>
__global__ __launch_bounds__(256) void k_adj_history(FrameDev f, const double* __restrict__ U,
                                                     const double* __restrict__ mu, double* __restrict__ hbs,
                                                     double* __restrict__ hbc, double* __restrict__ gs,
                                                     double* __restrict__ gc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long sec = (long long)blockIdx.x * 4 + wave;
  if (sec >= f.nsec) return;
  const int e = (int)(sec / f.nip), ip = (int)(sec % f.nip);
  double d[4], dm[4];
  section_deformation<D3>(f, sec, U, d);
  section_deformation<D3>(f, sec, mu, dm);
  // mu^T F_int = sum_ip wt L d_ip(mu) . s_ip (q = L sum_ip wt B_ip^T s_ip)
  const double wL = gl5_wt(ip) * 0.5 * f.geo[(long long)e * 10];
  const int def = f.secdef[e];
  const int f0 = f.def_off[def], nf = f.def_off[def + 1] - f0;
  const long long sb0 = f.steel_base[sec], cb0 = f.conc_base[sec];
  for (int i = lane; i < nf; i += 64) {
    const int fi = f0 + i;
    const int kind = f.fkind[fi];
    if (kind != 1 && kind != 2) continue;
    const double y = f.fy[fi];
    const double z = D3 ? f.fz[fi] : 0.0;
    const double strain = D3 ? d[0] - y * d[1] + z * d[2] : d[0] - y * d[1];
    const double vstrain = D3 ? dm[0] - y * dm[1] + z * dm[2] : dm[0] - y * dm[1];
    const double w = wL * f.fA[fi] * vstrain;
    double eb;
    if (kind == 1) {
      const long long n = f.nsteel, idx = sb0 + f.flocal[fi];
      double c[10], tb[10], cbo[10], thb[3];
#pragma unroll
      for (int k = 0; k < 10; k++) c[k] = f.st_c[k * n + idx];
#pragma unroll
      for (int k = 0; k < 9; k++) tb[k] = hbs[k * n + idx];
      tb[8] -= w;
      tb[9] = 0.0;
      steel02_vjp(f.steelp[f.fmat[fi]], c, f.kon_c[idx], strain, tb, cbo, thb, eb);
#pragma unroll
      for (int k = 0; k < 9; k++) hbs[k * n + idx] = cbo[k];
#pragma unroll
      for (int k = 0; k < 3; k++) gs[k * n + idx] += thb[k];
    } else {
      const long long n = f.nconc, idx = cb0 + f.flocal[fi];
      double c[5], tb[5], cbo[5], fcb;
#pragma unroll
      for (int k = 0; k < 5; k++) c[k] = f.cs_c[k * n + idx];
#pragma unroll
      for (int k = 0; k < 4; k++) tb[k] = hbc[k * n + idx];
      tb[3] -= w;
      tb[4] = 0.0;
      concrete02_vjp(f.concp[f.fmat[fi]], c, strain, tb, cbo, fcb, eb);
#pragma unroll
      for (int k = 0; k < 4; k++) hbc[k * n + idx] = cbo[k];
      gc[idx] += fcb;
    }
  }
}

// rhs = ubar + c2 vbar + c3 abar + g, with g given negated (as assembled)
__global__ void k_adj_rhs(int n, const double* __restrict__ ub, const double* __restrict__ vb,
                          const double* __restrict__ ab, const double* __restrict__ neg_g, double c2, double c3,
                          double* __restrict__ rhs) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) rhs[a] = ub[a] + c2 * vb[a] + c3 * ab[a] - neg_g[a];
}

__global__ void k_adj_newmark(int n, const double* __restrict__ M, const double* __restrict__ mu, AdjCoef k,
                              double* __restrict__ ub, double* __restrict__ vb, double* __restrict__ ab) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const double mmu = M[a] * mu[a];
  const double v = vb[a], w = ab[a];
  ub[a] = (k.c3 + k.alphaM * k.c2) * mmu - k.c2 * v - k.c3 * w;
  vb[a] = -(k.a3 + k.alphaM * k.a1) * mmu + k.a1 * v + k.a3 * w;
  ab[a] = -(k.a4 + k.alphaM * k.a2) * mmu + k.a2 * v + k.a4 * w;
}

// ubar[eq_i] += 2 U[dof_i] (distinct equations)
__global__ void k_loss_grad(int nl, const int* __restrict__ dofs, const int* __restrict__ eqs,
                            const double* __restrict__ U, double* __restrict__ ub) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nl && eqs[i] >= 0) ub[eqs[i]] += 2.0 * U[dofs[i]];
}

// acc[0] += sum_i U[dof_i]^2 (one workgroup, fixed reduction order)
__global__ __launch_bounds__(256) void k_loss_accum(int nl, const int* __restrict__ dofs,
                                                    const double* __restrict__ U, double* __restrict__ acc) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < nl; i += 256) {
    const double u = U[dofs[i]];
    s += u * u;
  }
  sh[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) acc[0] += sh[0];
}

// out[m * ncomp + c] = sum over the pool entries of material m of vals[c][.]
__global__ __launch_bounds__(256) void k_reduce_mat(int nmat, const int* __restrict__ ptr, const int* __restrict__ idx,
                                                    const double* __restrict__ vals, long long npool, int ncomp,
                                                    double* __restrict__ out) {
  const int m = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (m >= nmat) return;
  for (int c = 0; c < ncomp; c++) {
    double s = 0.0;
    for (int k = ptr[m] + lane; k < ptr[m + 1]; k += 64) s += vals[c * npool + idx[k]];
    s = wave_sum(s);
    if (lane == 0) out[(long long)m * ncomp + c] = s;
  }
}

void launch_adj_strain_adjoint(hipStream_t s, const FrameDev& f, bool d3, const double* U, const double* hbs,
                               const double* hbc, double* shat) {
  if (f.nsec <= 0) return;
  const dim3 g(ceil_div(f.nsec, 4)), b(256);
  if (d3) hipLaunchKernelGGL(k_adj_strain_adjoint<true>, g, b, 0, s, f, U, hbs, hbc, shat);
  else hipLaunchKernelGGL(k_adj_strain_adjoint<false>, g, b, 0, s, f, U, hbs, hbc, shat);
}
void launch_adj_history(hipStream_t s, const FrameDev& f, bool d3, const double* U, const double* mu, double* hbs,
                        double* hbc, double* gs, double* gc) {
  if (f.nsec <= 0) return;
  const dim3 g(ceil_div(f.nsec, 4)), b(256);
  if (d3) hipLaunchKernelGGL(k_adj_history<true>, g, b, 0, s, f, U, mu, hbs, hbc, gs, gc);
  else hipLaunchKernelGGL(k_adj_history<false>, g, b, 0, s, f, U, mu, hbs, hbc, gs, gc);
}
void launch_adj_rhs(hipStream_t s, int n, const double* ub, const double* vb, const double* ab, const double* neg_g,
                    double c2, double c3, double* rhs) {
  if (n > 0) hipLaunchKernelGGL(k_adj_rhs, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, ub, vb, ab, neg_g, c2, c3, rhs);
}
void launch_adj_newmark(hipStream_t s, int n, const double* M, const double* mu, const AdjCoef& k, double* ub,
                        double* vb, double* ab) {
  if (n > 0) hipLaunchKernelGGL(k_adj_newmark, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, M, mu, k, ub, vb, ab);
}
void launch_loss_grad(hipStream_t s, int nl, const int* dofs, const int* eqs, const double* U, double* ub) {
  if (nl > 0) hipLaunchKernelGGL(k_loss_grad, dim3(ceil_div(nl, 256)), dim3(256), 0, s, nl, dofs, eqs, U, ub);
}
void launch_loss_accum(hipStream_t s, int nl, const int* dofs, const double* U, double* acc) {
  if (nl > 0) hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(256), 0, s, nl, dofs, U, acc);
}
void launch_reduce_mat(hipStream_t s, int nmat, const int* ptr, const int* idx, const double* vals, long long npool,
                       int ncomp, double* out) {
  if (nmat > 0)
    hipLaunchKernelGGL(k_reduce_mat, dim3(ceil_div((long long)nmat * 64, 256)), dim3(256), 0, s, nmat, ptr, idx, vals,
                       npool, ncomp, out);
}

}  // namespace xg
