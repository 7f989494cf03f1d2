// stdBrick (8-node trilinear hexahedron, 2x2x2 Gauss) state determination
// on gfx950: update (nodal gather -> strain -> J2 return map), resisting
// force and tangent stiffness.  Follows Brick::update (Brick.cpp:719-835),
// Brick::formResidAndTangent (Brick.cpp:844-1024) and shp3d
// (SRC/interpolate/shp3d.cpp:32-172).
//
// Layout: material state is structure-of-arrays over points p = 8*e + g
// (g = 4i+2j+k, Brick.cpp:754-775); component c of array X lives at
// X[c*npts + p], so consecutive threads touch consecutive addresses.
#include "xg_j2.h"
#include "xg_kernels.h"

namespace xg {

// shp3d: derivatives w.r.t. x,y,z in shp[0..2][8], values in shp[3][8];
// returns the Jacobian determinant.  xl[j][a] = coordinate j of node a.
__device__ __forceinline__ double dev_shp3d(const double xn[3], double shp[4][8],
                                            const double (*xl)[8]) {
  const double ap1 = 1.0 + xn[0], am1 = 1.0 - xn[0];
  const double ap2 = 1.0 + xn[1], am2 = 1.0 - xn[1];
  const double ap3 = 1.0 + xn[2], am3 = 1.0 - xn[2];
  {
    double c1 = 0.125 * am1 * am2, c2 = 0.125 * am2 * am3, c3 = 0.125 * am1 * am3;
    shp[0][0] = -c2; shp[0][1] = c2; shp[1][0] = -c3; shp[1][3] = c3;
    shp[2][0] = -c1; shp[2][4] = c1; shp[3][0] = c1 * am3; shp[3][4] = c1 * ap3;
  }
  {
    double c1 = 0.125 * ap1 * ap2, c2 = 0.125 * ap2 * ap3, c3 = 0.125 * ap1 * ap3;
    shp[0][7] = -c2; shp[0][6] = c2; shp[1][5] = -c3; shp[1][6] = c3;
    shp[2][2] = -c1; shp[2][6] = c1; shp[3][2] = c1 * am3; shp[3][6] = c1 * ap3;
  }
  {
    double c1 = 0.125 * am1 * ap2, c2 = 0.125 * am2 * ap3, c3 = 0.125 * am1 * ap3;
    shp[0][4] = -c2; shp[0][5] = c2; shp[1][4] = -c3; shp[1][7] = c3;
    shp[2][3] = -c1; shp[2][7] = c1; shp[3][3] = c1 * am3; shp[3][7] = c1 * ap3;
  }
  {
    double c1 = 0.125 * ap1 * am2, c2 = 0.125 * ap2 * am3, c3 = 0.125 * ap1 * am3;
    shp[0][3] = -c2; shp[0][2] = c2; shp[1][1] = -c3; shp[1][2] = c3;
    shp[2][1] = -c1; shp[2][5] = c1; shp[3][1] = c1 * am3; shp[3][5] = c1 * ap3;
  }
  double xs[3][3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    xs[j][0] = (xl[j][1] - xl[j][0]) * shp[0][1] + (xl[j][2] - xl[j][3]) * shp[0][2] +
               (xl[j][5] - xl[j][4]) * shp[0][5] + (xl[j][6] - xl[j][7]) * shp[0][6];
    xs[j][1] = (xl[j][2] - xl[j][1]) * shp[1][2] + (xl[j][3] - xl[j][0]) * shp[1][3] +
               (xl[j][6] - xl[j][5]) * shp[1][6] + (xl[j][7] - xl[j][4]) * shp[1][7];
    xs[j][2] = (xl[j][4] - xl[j][0]) * shp[2][4] + (xl[j][5] - xl[j][1]) * shp[2][5] +
               (xl[j][6] - xl[j][2]) * shp[2][6] + (xl[j][7] - xl[j][3]) * shp[2][7];
  }
  double ad[3][3];
  ad[0][0] = xs[1][1] * xs[2][2] - xs[1][2] * xs[2][1];
  ad[0][1] = xs[2][1] * xs[0][2] - xs[2][2] * xs[0][1];
  ad[0][2] = xs[0][1] * xs[1][2] - xs[0][2] * xs[1][1];
  ad[1][0] = xs[1][2] * xs[2][0] - xs[1][0] * xs[2][2];
  ad[1][1] = xs[2][2] * xs[0][0] - xs[2][0] * xs[0][2];
  ad[1][2] = xs[0][2] * xs[1][0] - xs[0][0] * xs[1][2];
  ad[2][0] = xs[1][0] * xs[2][1] - xs[1][1] * xs[2][0];
  ad[2][1] = xs[2][0] * xs[0][1] - xs[2][1] * xs[0][0];
  ad[2][2] = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
  const double xsj = xs[0][0] * ad[0][0] + xs[0][1] * ad[1][0] + xs[0][2] * ad[2][0];
  const double rxsj = 1.0 / xsj;
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) xs[i][j] = ad[i][j] * rxsj;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    double c1 = shp[0][k] * xs[0][0] + shp[1][k] * xs[1][0] + shp[2][k] * xs[2][0];
    double c2 = shp[0][k] * xs[0][1] + shp[1][k] * xs[1][1] + shp[2][k] * xs[2][1];
    double c3 = shp[0][k] * xs[0][2] + shp[1][k] * xs[1][2] + shp[2][k] * xs[2][2];
    shp[0][k] = c1;
    shp[1][k] = c2;
    shp[2][k] = c3;
  }
  return xsj;
}

__device__ __forceinline__ void brick_gauss(int g, double xn[3]) {
  const double s = 1.0 / sqrt(3.0);  // Brick::one_over_root3
  xn[0] = ((g >> 2) & 1) ? s : -s;
  xn[1] = ((g >> 1) & 1) ? s : -s;
  xn[2] = (g & 1) ? s : -s;
}

constexpr int kBrickUpdElems = 32;  // elements per 256-thread block

// One thread per Gauss point.  The block's 32 elements gather their 8 nodal
// coordinates and trial displacements into LDS once (thread (e,g) loads node
// g of element e), then every point evaluates shp3d, the strain of
// Brick::update and the J2 return map, and writes its state, stress and the
// two tangent coefficients (tc[0] = c2, tc[1] = c3).
__global__ __launch_bounds__(256) void k_brick_update(
    int nele, const int* __restrict__ conn, const int* __restrict__ mat,
    const J2Param* __restrict__ params, const double* __restrict__ crd,
    const double* __restrict__ U, int ndf, double dt, long long npts,
    const double* __restrict__ epn, const double* __restrict__ xin, double* __restrict__ ept,
    double* __restrict__ xit, double* __restrict__ sig, double* __restrict__ tc,
    int* __restrict__ nfail) {
  __shared__ double s_x[kBrickUpdElems][3][8];
  __shared__ double s_u[kBrickUpdElems][8][3];
  const int t = threadIdx.x;
  const int le = t >> 3, g = t & 7;
  const long long e = (long long)blockIdx.x * kBrickUpdElems + le;
  const bool active = e < nele;
  if (active) {
    const int node = conn[e * 8 + g];
    s_x[le][0][g] = crd[3LL * node + 0];
    s_x[le][1][g] = crd[3LL * node + 1];
    s_x[le][2][g] = crd[3LL * node + 2];
    s_u[le][g][0] = U[(long long)ndf * node + 0];
    s_u[le][g][1] = U[(long long)ndf * node + 1];
    s_u[le][g][2] = U[(long long)ndf * node + 2];
  }
  __syncthreads();
  if (!active) return;
  const long long p = e * 8 + g;
  double xn[3], shp[4][8];
  brick_gauss(g, xn);
  dev_shp3d(xn, shp, s_x[le]);
  double s[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double ul0 = s_u[le][j][0], ul1 = s_u[le][j][1], ul2 = s_u[le][j][2];
    s[0] += shp[0][j] * ul0;
    s[1] += shp[1][j] * ul1;
    s[2] += shp[2][j] * ul2;
    s[3] += shp[1][j] * ul0 + shp[0][j] * ul1;
    s[4] += shp[2][j] * ul1 + shp[1][j] * ul2;
    s[5] += shp[2][j] * ul0 + shp[0][j] * ul2;
  }
  const double eps[6] = {s[0], s[1], s[2], 0.50 * s[3], 0.50 * s[4], 0.50 * s[5]};
  double en[6];
#pragma unroll
  for (int c = 0; c < 6; c++) en[c] = epn[c * npts + p];
  const double xn0 = xin[p];
  const J2Param prm = params[mat[e]];
  J2Out o;
  if (!j2_return_map(prm, dt, eps, en, xn0, o)) {
    nfail[0] = 1;  // flag only; the point's trial state is left untouched
    return;
  }
  // streaming (non-temporal) stores: this kernel never reads them back, and
  // keeping them out of the caches' way measured 25% faster at 64^3
#pragma unroll
  for (int c = 0; c < 6; c++) __builtin_nontemporal_store(o.ep[c], &ept[c * npts + p]);
  __builtin_nontemporal_store(o.xi, &xit[p]);
#pragma unroll
  for (int c = 0; c < 6; c++) __builtin_nontemporal_store(o.sig[c], &sig[c * npts + p]);
  __builtin_nontemporal_store(o.c2, &tc[p]);
  __builtin_nontemporal_store(o.c3, &tc[npts + p]);
}

// Resisting force of each element, one thread per element, Gauss points in
// the reference order (Brick.cpp:925-1018 with tang_flag = 0):
//   R += B^T (sigma dvol) - dvol b N.
__global__ __launch_bounds__(128) void k_brick_resid(
    int nele, const int* __restrict__ conn, const double* __restrict__ crd,
    const double* __restrict__ bforce, long long npts, const double* __restrict__ sig,
    double* __restrict__ Re) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nele) return;
  double xl[3][8];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int node = conn[e * 8 + a];
    xl[0][a] = crd[3LL * node + 0];
    xl[1][a] = crd[3LL * node + 1];
    xl[2][a] = crd[3LL * node + 2];
  }
  double b[3] = {0, 0, 0};
  if (bforce) {
    b[0] = bforce[3 * e + 0];
    b[1] = bforce[3 * e + 1];
    b[2] = bforce[3 * e + 2];
  }
  double r[24];
#pragma unroll
  for (int i = 0; i < 24; i++) r[i] = 0.0;
  for (int g = 0; g < 8; g++) {
    double xn[3], shp[4][8];
    brick_gauss(g, xn);
    const double dvol = 1.0 * dev_shp3d(xn, shp, xl);
    const long long p = e * 8 + g;
    double st[6];
#pragma unroll
    for (int c = 0; c < 6; c++) st[c] = sig[c * npts + p] * dvol;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const double r0 = shp[0][j] * st[0] + shp[1][j] * st[3] + shp[2][j] * st[5];
      const double r1 = shp[1][j] * st[1] + shp[0][j] * st[3] + shp[2][j] * st[4];
      const double r2 = shp[2][j] * st[2] + shp[1][j] * st[4] + shp[0][j] * st[5];
      r[3 * j + 0] += r0;
      r[3 * j + 0] -= dvol * b[0] * shp[3][j];
      r[3 * j + 1] += r1;
      r[3 * j + 1] -= dvol * b[1] * shp[3][j];
      r[3 * j + 2] += r2;
      r[3 * j + 2] -= dvol * b[2] * shp[3][j];
    }
  }
#pragma unroll
  for (int i = 0; i < 24; i++) Re[e * 24 + i] = r[i];
}

// Block (J, K) of the element tangent, sum over the 8 Gauss points of
// BJ^T (D dvol) BK (Brick::formResidAndTangent, Brick.cpp:970-1004), from
// the points' tangents D (upper triangle), shape derivatives shp[g][dir][node]
// and volume weights dv staged by the caller.
__device__ __forceinline__ void brick_block(int J, int K, const double (*D)[21], const double (*shp)[3][8],
                                            const double* dv, double k9[3][3]) {
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) k9[p][q] = 0.0;
  for (int g = 0; g < 8; g++) {
    const double dvol = dv[g];
    double dd[6][6];
    {
      int idx = 0;
#pragma unroll
      for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = a; b < 6; b++) {
          const double v = D[g][idx++] * dvol;
          dd[a][b] = v;
          dd[b][a] = v;
        }
    }
    const double jx = shp[g][0][J], jy = shp[g][1][J], jz = shp[g][2][J];
    const double kx = shp[g][0][K], ky = shp[g][1][K], kz = shp[g][2][K];
    // BJ^T rows (computeB, Brick.cpp:1047-1076):
    //   p=0: [jx 0 0 jy 0 jz], p=1: [0 jy 0 jx jz 0], p=2: [0 0 jz 0 jy jx]
    double bt[3][6] = {{jx, 0, 0, jy, 0, jz}, {0, jy, 0, jx, jz, 0}, {0, 0, jz, 0, jy, jx}};
    double btd[3][6];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 6; q++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 6; k++) acc += bt[p][k] * dd[k][q];
        btd[p][q] = acc;
      }
    // BK columns: q=0: [kx 0 0 ky 0 kz], q=1: [0 ky 0 kx kz 0], q=2: [0 0 kz 0 ky kx]
    const double bk[6][3] = {{kx, 0, 0}, {0, ky, 0}, {0, 0, kz}, {ky, kx, 0}, {0, kz, ky}, {kz, 0, kx}};
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 6; k++) acc += btd[p][k] * bk[k][q];
        k9[p][q] += acc;
      }
  }
}

constexpr int kBrickTanElems = 4;  // one 64-lane wavefront per element

// Tangent stiffness K_e = sum_g B^T (D dvol) B (Brick.cpp:925-1018,
// tang_flag = 1).  One wavefront per element; lane (J,K) forms the 3x3
// block of node pair (J,K) summing the eight Gauss points in order.  Shape
// derivatives, dvol and the point tangents (re-formed from each point's
// stress and coefficients by lanes 8..15) are staged in LDS.
__global__ __launch_bounds__(256) void k_brick_tangent(
    int nele, const int* __restrict__ conn, const int* __restrict__ mat,
    const J2Param* __restrict__ params, const double* __restrict__ crd, long long npts,
    const double* __restrict__ sig, const double* __restrict__ tc, double* __restrict__ Ke) {
  __shared__ double s_xl[kBrickTanElems][3][8];
  __shared__ double s_shp[kBrickTanElems][8][3][8];  // [g][dir][node]
  __shared__ double s_dv[kBrickTanElems][8];
  __shared__ double s_D[kBrickTanElems][8][21];
  const int t = threadIdx.x;
  const int le = t >> 6, lane = t & 63;
  const long long e = (long long)blockIdx.x * kBrickTanElems + le;
  const bool active = e < nele;
  if (active) {
    if (lane < 24) {
      const int node_slot = lane & 7, dir = lane >> 3;
      const int node = conn[e * 8 + node_slot];
      s_xl[le][dir][node_slot] = crd[3LL * node + dir];
    }
    if (lane >= 32 && lane < 40) {
      const int g = lane - 32;
      const long long p = e * 8 + g;
      double st[6], dg[21];
#pragma unroll
      for (int c = 0; c < 6; c++) st[c] = sig[c * npts + p];
      j2_tangent_from_stress(params[mat[e]], st, tc[p], tc[npts + p], dg);
#pragma unroll
      for (int c = 0; c < 21; c++) s_D[le][g][c] = dg[c];
    }
  }
  __syncthreads();
  if (active && lane < 8) {
    double xn[3], shp[4][8];
    brick_gauss(lane, xn);
    const double dvol = 1.0 * dev_shp3d(xn, shp, s_xl[le]);
    s_dv[le][lane] = dvol;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int a = 0; a < 8; a++) s_shp[le][lane][d][a] = shp[d][a];
  }
  __syncthreads();
  if (!active) return;
  const int J = lane >> 3, K = lane & 7;
  double k9[3][3];
  brick_block(J, K, s_D[le], s_shp[le], s_dv[le], k9);
  double* out = Ke + e * 576;
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) out[(3 * J + p) * 24 + 3 * K + q] = k9[p][q];
}

// formTangent of a stdBrick-only model straight into the CSR values, without
// element matrices: one wavefront per node block row I (equations 3I ..
// 3I+2, every node's DOFs all free or all fixed).  The elements around the
// node are taken eight at a time in ascending tag order; lane (s, g) stages
// element s's shape derivatives at Gauss point g, the point's stress
// normal n and its tangent coefficients in LDS, then lane (s, b) forms the
// element's block (a, b), a the node's local index, and puts it at its
// neighbour's position; each CSR value then sums its elements in tag order,
// the order of the element-by-element assembly (exact zeros stand for the
// elements that do not touch it).  The block BaT D Bb of the J2 tangent
// D = K 1(x)1 + 2G Idev + c2 n(x)n + c3 (Idev - n(x)n) (J2Plasticity.cpp:
// 361-391) is formed in closed form,
//   lam ga gb^T + mu ((ga . gb) I + gb ga^T) + gam (N ga)(N gb)^T,
// lam = K - (2G + c3)/3, mu = G + c3/2, gam = c2 - c3, ga the shape
// derivatives of node a, N the normal as a symmetric 3x3 tensor: 57 FMAs per
// point instead of the 324 of the product with the 6x6 matrix.  Table (per
// block row I, entries fz_ptr[I] .. fz_ptr[I+1]): the element, the node's
// local index in it, and for each of its 8 nodes the neighbour position in
// row I (255: a fixed node).
constexpr int kFzMaxNb = 32;  // neighbour blocks per row the LDS holds
constexpr int kFzPt = 33;     // staged doubles per (element, point): shp 24, coefficients 3, n 6
__global__ __launch_bounds__(64) void k_brick_tangent_csr(
    int nbr, const int* __restrict__ rowptr, const int* __restrict__ fz_ptr, const int* __restrict__ fz_ele,
    const unsigned char* __restrict__ fz_a, const unsigned char* __restrict__ fz_tgt,
    const int* __restrict__ conn, const int* __restrict__ mat, const J2Param* __restrict__ params,
    const double* __restrict__ crd, long long npts, const double* __restrict__ sig,
    const double* __restrict__ tc, double* __restrict__ vals) {
  // staging [s][g][kFzPt]; reused for the blocks by neighbour, acc[k][s][9]
  __shared__ double lds[kFzMaxNb * 8 * 9];
  static_assert(8 * 8 * kFzPt <= kFzMaxNb * 8 * 9, "staging fits the block buffer");
  const int lane = threadIdx.x, s = lane >> 3, g = lane & 7;
  // workgroups are dealt to the 8 XCDs round-robin: give each XCD one
  // contiguous range of rows, so that the rows sharing an element are
  // mostly formed under one L2
  const int per = (nbr + 7) / 8;
  const int I0 = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int I1 = min(nbr, ((int)blockIdx.x & 7) * per + per);
  for (int I = I0; I < I1; I += (int)gridDim.x >> 3) {
    const int b0 = rowptr[3 * I], len = rowptr[3 * I + 1] - b0, nnb = len / 3;
    const int p0 = fz_ptr[I], ne = fz_ptr[I + 1] - p0;
    double run[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // output o = lane + 64 i, o = p * len + k
    for (int pass = 0; pass * 8 < ne; pass++) {
      const int es = pass * 8 + s;
      const bool vs = es < ne;
      const int e = vs ? fz_ele[p0 + es] : 0;
      if (vs) {
        double xl[3][8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int nd = conn[8LL * e + k];
#pragma unroll
          for (int d = 0; d < 3; d++) xl[d][k] = crd[3LL * nd + d];
        }
        const long long pt = 8LL * e + g;
        double st[6], nn[6];
#pragma unroll
        for (int c = 0; c < 6; c++) st[c] = sig[c * npts + pt];
        const double c2 = tc[pt], c3 = tc[npts + pt];
        const J2Param pr = params[mat[e]];
        double xn[3], shp[4][8];
        brick_gauss(g, xn);
        const double dvol = 1.0 * dev_shp3d(xn, shp, xl);
        j2_stress_normal(st, nn);
        double* o = lds + (s * 8 + g) * kFzPt;
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int k = 0; k < 8; k++) o[d * 8 + k] = shp[d][k];
        o[24] = (pr.K - (2.0 * pr.G + c3) * (1.0 / 3.0)) * dvol;
        o[25] = (pr.G + 0.5 * c3) * dvol;
        o[26] = (c2 - c3) * dvol;
#pragma unroll
        for (int c = 0; c < 6; c++) o[27 + c] = nn[c];
      }
      __syncthreads();
      double k9[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (vs) {
        const int a = fz_a[p0 + es];
        for (int gg = 0; gg < 8; gg++) {
          const double* q = lds + (s * 8 + gg) * kFzPt;
          const double ga[3] = {q[a], q[8 + a], q[16 + a]};
          const double gb[3] = {q[g], q[8 + g], q[16 + g]};
          const double lam = q[24], mu = q[25], gam = q[26];
          // N rows: (n00 n01 n20), (n01 n11 n12), (n20 n12 n22)
          const double n00 = q[27], n11 = q[28], n22 = q[29], n01 = q[30], n12 = q[31], n20 = q[32];
          const double va[3] = {n00 * ga[0] + n01 * ga[1] + n20 * ga[2], n01 * ga[0] + n11 * ga[1] + n12 * ga[2],
                                n20 * ga[0] + n12 * ga[1] + n22 * ga[2]};
          const double vb[3] = {n00 * gb[0] + n01 * gb[1] + n20 * gb[2], n01 * gb[0] + n11 * gb[1] + n12 * gb[2],
                                n20 * gb[0] + n12 * gb[1] + n22 * gb[2]};
          const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
          for (int p = 0; p < 3; p++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
              double v = lam * ga[p] * gb[r] + mu * gb[p] * ga[r] + gam * va[p] * vb[r];
              if (p == r) v += mu * dot;
              k9[p][r] += v;
            }
        }
      }
      __syncthreads();
      for (int i = lane; i < nnb * 72; i += 64) lds[i] = 0.0;
      __syncthreads();
      const int tg = vs ? fz_tgt[8LL * (p0 + es) + g] : 255;
      if (tg != 255) {
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
          for (int q = 0; q < 3; q++) lds[(tg * 8 + s) * 9 + 3 * p + q] = k9[p][q];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const int o = lane + 64 * i;
        if (o < 3 * len) {
          const int p = o / len, k = o - p * len, j = k / 3, q = k - 3 * j;
          double r = run[i];
#pragma unroll
          for (int s2 = 0; s2 < 8; s2++) r += lds[(j * 8 + s2) * 9 + 3 * p + q];
          run[i] = r;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const int o = lane + 64 * i;
      if (o < 3 * len) vals[b0 + o] = run[i];
    }
  }
}

// commitState of every point (J2Plasticity.cpp:516-522): n <- n+1.
__global__ void k_j2_commit(long long npts, int ncomp, const double* __restrict__ ept,
                            const double* __restrict__ xit, double* __restrict__ epn,
                            double* __restrict__ xin) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  for (int c = 0; c < ncomp; c++) epn[c * npts + p] = ept[c * npts + p];
  xin[p] = xit[p];
}

void launch_brick_update(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                         const double* crd, const double* U, int ndf, double dt, long long npts,
                         const double* epn, const double* xin, double* ept, double* xit, double* sig,
                         double* tc, int* nfail) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_brick_update, dim3(ceil_div(nele, kBrickUpdElems)), dim3(256), 0, s, nele,
                     conn, mat, prm, crd, U, ndf, dt, npts, epn, xin, ept, xit, sig, tc, nfail);
}

void launch_brick_resid(hipStream_t s, int nele, const int* conn, const double* crd,
                        const double* bforce, long long npts, const double* sig, double* Re) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_brick_resid, dim3(ceil_div(nele, 128)), dim3(128), 0, s, nele, conn, crd,
                     bforce, npts, sig, Re);
}

void launch_brick_tangent(hipStream_t s, int nele, const int* conn, const int* mat, const J2Param* prm,
                          const double* crd, long long npts, const double* sig, const double* tc, double* Ke) {
  if (nele <= 0) return;
  hipLaunchKernelGGL(k_brick_tangent, dim3(ceil_div(nele, kBrickTanElems)), dim3(256), 0, s, nele,
                     conn, mat, prm, crd, npts, sig, tc, Ke);
}

void launch_brick_tangent_csr(hipStream_t s, int nbr, const int* rowptr, const int* fz_ptr, const int* fz_ele,
                              const unsigned char* fz_a, const unsigned char* fz_tgt, const int* conn,
                              const int* mat, const J2Param* prm, const double* crd, long long npts,
                              const double* sig, const double* tc, double* vals) {
  if (nbr <= 0) return;
  const int grid = 8 * std::min((nbr + 7) / 8, 1 << 17);  // a multiple of the 8 XCDs
  hipLaunchKernelGGL(k_brick_tangent_csr, dim3(grid), dim3(64), 0, s, nbr, rowptr, fz_ptr, fz_ele, fz_a, fz_tgt,
                     conn, mat, prm, crd, npts, sig, tc, vals);
}

void launch_j2_commit(hipStream_t s, long long npts, int ncomp, const double* ept, const double* xit,
                      double* epn, double* xin) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(k_j2_commit, dim3(ceil_div(npts, 256)), dim3(256), 0, s, npts, ncomp, ept,
                     xit, epn, xin);
}

}  // namespace xg
