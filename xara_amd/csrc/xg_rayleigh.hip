// Rayleigh stiffness-proportional damping of the element groups
// (Element::getDamp and Element::getRayleighDampingForces,
// Element.cpp:181-200 and 283-323; Newmark::formEleTangent,
// Newmark.cpp:262-289).  The element matrices are the groups' K_e blocks,
// row-major nd x nd per element, elements in group order; Kc holds the
// tangent of the last commit in the same layout.
#include <algorithm>

#include "xg_common.h"

namespace xg {

// K_e <- K_e + c2 (betaK K_e + betaKc Kc) entry by entry: the current
// tangent with the stiffness-proportional part of c2 C added in the
// reference's order (getDamp sums betaK K then betaKc Kc; formEleTangent
// adds c1 K, c1 = 1, then c2 C).  Mass-proportional terms are added on the
// diagonal by the assembly of the nodal masses.
__global__ void __launch_bounds__(256) k_ke_damp(long long n, double c2, double betaK, double betaKc,
                                                 const double* __restrict__ Kc, double* __restrict__ Ke) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
    const double k = Ke[i];
    double c = 0.0;
    if (betaK != 0.0) c = betaK * k;
    if (betaKc != 0.0) c = c + betaKc * Kc[i];
    Ke[i] = k + c2 * c;
  }
}

// Re_e += (betaK Kt_e + betaKc Kc_e) v_e: one thread per element row, v_e
// gathered from the nodal trial velocities (node-major, ndf per node).
__global__ void __launch_bounds__(256) k_rayleigh_force(int nele, int nen, int ndf, const int* __restrict__ conn,
                                                        const double* __restrict__ Kt, const double* __restrict__ Kc,
                                                        double betaK, double betaKc, const double* __restrict__ V,
                                                        double* __restrict__ Re) {
  const int nd = nen * ndf;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)nele * nd) return;
  const int e = (int)(t / nd), i = (int)(t - (long long)e * nd);
  const double* kt = Kt + ((long long)e * nd + i) * nd;
  const double* kc = Kc ? Kc + ((long long)e * nd + i) * nd : nullptr;
  const int* en = conn + (long long)e * nen;
  double f = 0.0;
  for (int a = 0; a < nen; a++) {
    const double* v = V + (long long)en[a] * ndf;
    for (int j = 0; j < ndf; j++) {
      const int c = a * ndf + j;
      double d = 0.0;
      if (betaK != 0.0) d = betaK * kt[c];
      if (betaKc != 0.0) d = d + betaKc * kc[c];
      f += d * v[j];
    }
  }
  Re[t] += f;
}

void launch_ke_damp(hipStream_t s, long long n, double c2, double betaK, double betaKc, const double* Kc,
                    double* Ke) {
  if (n <= 0) return;
  const int grid = (int)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ke_damp, dim3(grid), dim3(256), 0, s, n, c2, betaK, betaKc, Kc, Ke);
}

void launch_rayleigh_force(hipStream_t s, int nele, int nen, int ndf, const int* conn, const double* Kt,
                           const double* Kc, double betaK, double betaKc, const double* V, double* Re) {
  const long long n = (long long)nele * nen * ndf;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rayleigh_force, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nele, nen, ndf, conn,
                     Kt, Kc, betaK, betaKc, V, Re);
}

}  // namespace xg
