// FourNodeQuad (bilinear, 2x2 Gauss, plane strain J2) state determination
// on gfx950.  Follows FourNodeQuad::update / getTangentStiff /
// getResistingForce / shapeFunction (FourNodeQuad.cpp:190-281, :507-553,
// :1128-1195) with GaussLegendre<2,4> (GaussLegendre2D.hpp:23-38) and
// J2PlaneStrain (J2PlaneStrain.cpp:124-204).
//
// Plane-strain J2 state per point (SoA, p = 4e + i): plastic strain tensor
// components 00,11,22,01 and xi; stress 00,11,01; tangent upper triangle
// (0,0)(0,1)(0,2)(1,1)(1,2)(2,2).
#include "xg_j2.h"
#include "xg_kernels.h"

namespace xg {

__device__ __forceinline__ void quad_gauss(int i, double& xi, double& eta) {
  const double a = 0.57735026918962573;
  xi = (i & 2) ? a : -a;   // (-,-),(-,+),(+,-),(+,+)
  eta = (i & 1) ? a : -a;
}

// shapeFunction: shp[0..1][4] derivatives, shp[2][4] values; returns detJ.
__device__ __forceinline__ double quad_shape(double xi, double eta, const double (*c)[2], double shp[3][4]) {
  const double oneMinuseta = 1.0 - eta, onePluseta = 1.0 + eta;
  const double oneMinusxi = 1.0 - xi, onePlusxi = 1.0 + xi;
  shp[2][0] = 0.25 * oneMinusxi * oneMinuseta;
  shp[2][1] = 0.25 * onePlusxi * oneMinuseta;
  shp[2][2] = 0.25 * onePlusxi * onePluseta;
  shp[2][3] = 0.25 * oneMinusxi * onePluseta;
  double J00 = 0.25 * (-c[0][0] * oneMinuseta + c[1][0] * oneMinuseta + c[2][0] * (onePluseta) -
                       c[3][0] * (onePluseta));
  double J01 = 0.25 * (-c[0][0] * oneMinusxi - c[1][0] * onePlusxi + c[2][0] * onePlusxi +
                       c[3][0] * oneMinusxi);
  double J10 = 0.25 * (-c[0][1] * oneMinuseta + c[1][1] * oneMinuseta + c[2][1] * onePluseta -
                       c[3][1] * onePluseta);
  double J11 = 0.25 * (-c[0][1] * oneMinusxi - c[1][1] * onePlusxi + c[2][1] * onePlusxi +
                       c[3][1] * oneMinusxi);
  const double detJ = J00 * J11 - J01 * J10;
  const double oneOverdetJ = 1.0 / detJ;
  const double L00 = 0.25 * (J11 * oneOverdetJ);
  const double L10 = 0.25 * (-J01 * oneOverdetJ);
  const double L01 = 0.25 * (-J10 * oneOverdetJ);
  const double L11 = 0.25 * (J00 * oneOverdetJ);
  const double L00m = L00 * oneMinuseta, L00p = L00 * onePluseta;
  const double L01m = L01 * oneMinusxi, L01p = L01 * onePlusxi;
  const double L10m = L10 * oneMinuseta, L10p = L10 * onePluseta;
  const double L11m = L11 * oneMinusxi, L11p = L11 * onePlusxi;
  shp[0][0] = -L00m - L01m;
  shp[0][1] = L00m - L01p;
  shp[0][2] = L00p + L01p;
  shp[0][3] = -L00p + L01m;
  shp[1][0] = -L10m - L11m;
  shp[1][1] = L10m - L11p;
  shp[1][2] = L10p + L11p;
  shp[1][3] = -L10p + L11m;
  return detJ;
}

constexpr int kQuadUpdElems = 64;

__global__ __launch_bounds__(256) void k_quad_update(
    int nele, const int* __restrict__ conn, const int* __restrict__ mat,
    const J2Param* __restrict__ params, const double* __restrict__ crd,
    const double* __restrict__ U, int ndf, double dt, long long npts,
    const double* __restrict__ epn, const double* __restrict__ xin, double* __restrict__ ept,
    double* __restrict__ xit, double* __restrict__ sig, double* __restrict__ D,
    int* __restrict__ nfail) {
  __shared__ double s_c[kQuadUpdElems][4][2];
  __shared__ double s_u[kQuadUpdElems][4][2];
  const int t = threadIdx.x;
  const int le = t >> 2, i = t & 3;
  const long long e = (long long)blockIdx.x * kQuadUpdElems + le;
  const bool active = e < nele;
  if (active) {
    const int node = conn[e * 4 + i];
    s_c[le][i][0] = crd[2LL * node + 0];
    s_c[le][i][1] = crd[2LL * node + 1];
    s_u[le][i][0] = U[(long long)ndf * node + 0];
    s_u[le][i][1] = U[(long long)ndf * node + 1];
  }
  __syncthreads();
  if (!active) return;
  const long long p = e * 4 + i;
  double xi, eta, shp[3][4];
  quad_gauss(i, xi, eta);
  quad_shape(xi, eta, s_c[le], shp);
  double e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    e0 += shp[0][b] * s_u[le][b][0];
    e1 += shp[1][b] * s_u[le][b][1];
    e2 += shp[0][b] * s_u[le][b][1] + shp[1][b] * s_u[le][b][0];
  }
  const double eps[6] = {e0, e1, 0.0, 0.50 * e2, 0.0, 0.0};
  const double en[6] = {epn[p], epn[npts + p], epn[2 * npts + p], epn[3 * npts + p], 0.0, 0.0};
  const J2Param prm = params[mat[e]];
  J2Out o;
  if (!j2_return_map(prm, dt, eps, en, xin[p], o)) {
    nfail[0] = 1;
    return;
  }
  ept[p] = o.ep[0];
  ept[npts + p] = o.ep[1];
  ept[2 * npts + p] = o.ep[2];
  ept[3 * npts + p] = o.ep[3];
  xit[p] = o.xi;
  sig[p] = o.sig[0];
  sig[npts + p] = o.sig[1];
  sig[2 * npts + p] = o.sig[3];
  double D6[21];
  j2_tangent_upper(prm, o.c2, o.c3, o.n, D6);
  D[p] = D6[tri_index(6, 0, 0)];
  D[npts + p] = D6[tri_index(6, 0, 1)];
  D[2 * npts + p] = D6[tri_index(6, 0, 3)];
  D[3 * npts + p] = D6[tri_index(6, 1, 1)];
  D[4 * npts + p] = D6[tri_index(6, 1, 3)];
  D[5 * npts + p] = D6[tri_index(6, 3, 3)];
}

__device__ __forceinline__ void quad_load_coords(long long e, const int* conn, const double* crd,
                                                 double c[4][2]) {
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int node = conn[e * 4 + a];
    c[a][0] = crd[2LL * node + 0];
    c[a][1] = crd[2LL * node + 1];
  }
}

__global__ void k_quad_resid(int nele, const int* __restrict__ conn, const double* __restrict__ crd,
                             const double* __restrict__ thk, const double* __restrict__ bforce,
                             long long npts, const double* __restrict__ sig, double* __restrict__ Re) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  double c[4][2];
  quad_load_coords(e, conn, crd, c);
  const double b0 = bforce ? bforce[2 * e] : 0.0, b1 = bforce ? bforce[2 * e + 1] : 0.0;
  double P[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    double xi, eta, shp[3][4];
    quad_gauss(i, xi, eta);
    double dvol = quad_shape(xi, eta, c, shp);
    dvol *= (thk[e] * 1.0);
    const long long p = e * 4 + i;
    const double s0 = sig[p], s1 = sig[npts + p], s2 = sig[2 * npts + p];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      P[2 * a] += dvol * (shp[0][a] * s0 + shp[1][a] * s2);
      P[2 * a + 1] += dvol * (shp[1][a] * s1 + shp[0][a] * s2);
      P[2 * a] -= dvol * (shp[2][a] * b0);
      P[2 * a + 1] -= dvol * (shp[2][a] * b1);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) Re[e * 8 + k] = P[k];
}

__global__ void k_quad_tangent(int nele, const int* __restrict__ conn, const double* __restrict__ crd,
                               const double* __restrict__ thk, long long npts,
                               const double* __restrict__ D, double* __restrict__ Ke) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  double c[4][2];
  quad_load_coords(e, conn, crd, c);
  double K[64];
#pragma unroll
  for (int k = 0; k < 64; k++) K[k] = 0.0;
  for (int i = 0; i < 4; i++) {
    double xi, eta, shp[3][4];
    quad_gauss(i, xi, eta);
    double dvol = quad_shape(xi, eta, c, shp);
    dvol *= (thk[e] * 1.0);
    const long long p = e * 4 + i;
    const double D00 = D[p], D01 = D[npts + p], D02 = D[2 * npts + p];
    const double D11 = D[3 * npts + p], D12 = D[4 * npts + p], D22 = D[5 * npts + p];
    const double D10 = D01, D20 = D02, D21 = D12;
#pragma unroll
    for (int al = 0; al < 4; al++) {
#pragma unroll
      for (int be = 0; be < 4; be++) {
        const double DB00 = dvol * (D00 * shp[0][be] + D02 * shp[1][be]);
        const double DB10 = dvol * (D10 * shp[0][be] + D12 * shp[1][be]);
        const double DB20 = dvol * (D20 * shp[0][be] + D22 * shp[1][be]);
        const double DB01 = dvol * (D01 * shp[1][be] + D02 * shp[0][be]);
        const double DB11 = dvol * (D11 * shp[1][be] + D12 * shp[0][be]);
        const double DB21 = dvol * (D21 * shp[1][be] + D22 * shp[0][be]);
        const int ia = 2 * al, ib = 2 * be;
        K[ia * 8 + ib] += shp[0][al] * DB00 + shp[1][al] * DB20;
        K[ia * 8 + ib + 1] += shp[0][al] * DB01 + shp[1][al] * DB21;
        K[(ia + 1) * 8 + ib] += shp[1][al] * DB10 + shp[0][al] * DB20;
        K[(ia + 1) * 8 + ib + 1] += shp[1][al] * DB11 + shp[0][al] * DB21;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 64; k++) Ke[e * 64 + k] = K[k];
}

void launch_quad_update(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                        const double* crd, const double* U, int ndf, double dt, long long npts,
                        const double* epn, const double* xin, double* ept, double* xit, double* sig,
                        double* D, int* nfail) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_quad_update, dim3(ceil_div(nele, kQuadUpdElems)), dim3(256), 0, s, nele, conn,
                     mat, prm, crd, U, ndf, dt, npts, epn, xin, ept, xit, sig, D, nfail);
}
void launch_quad_resid(hipStream_t s, int nele, const int* conn, const double* crd, const double* thk,
                       const double* bforce, long long npts, const double* sig, double* Re) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_quad_resid, dim3(ceil_div(nele, 128)), dim3(128), 0, s, nele, conn, crd, thk,
                     bforce, npts, sig, Re);
}
void launch_quad_tangent(hipStream_t s, int nele, const int* conn, const double* crd, const double* thk,
                         long long npts, const double* D, double* Ke) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_quad_tangent, dim3(ceil_div(nele, 128)), dim3(128), 0, s, nele, conn, crd, thk,
                     npts, D, Ke);
}

}  // namespace xg
