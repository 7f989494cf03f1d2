// J2 plasticity material-point update for gfx950: backward-Euler radial
// return with linear + saturation isotropic hardening and optional linear
// viscosity, following J2Plasticity::plastic_integrator
// (SRC/material/plastic/J2Plasticity.cpp:228-394) in symmetric storage.
//
// Strain, plastic strain and stress are symmetric tensors held as six tensor
// components in the order 00,11,22,01,12,20 (J2ThreeDimensional.cpp:189-203);
// the consistent tangent is the Voigt 6x6 matrix of J2ThreeDimensional::
// getTangent (:205-234) -- symmetric, so only its upper triangle is formed.
//
// The tangent of plastic_integrator (J2Plasticity.cpp:361-391) is
//   K I(x)I + 2G Idev + c2 n(x)n + c3 (Idev - n(x)n)
// with two scalars c2, c3 (zero at an elastic point) and the unit normal n
// of the trial deviatoric stress.  The radial return keeps the deviatoric
// stress parallel to n, so a point stores c2 and c3 only (16 bytes instead
// of the 168 of the upper triangle) and the tangent is re-formed from the
// stress where it is used.
#pragma once
#include "xg_common.h"

namespace xg {

struct J2Out {
  double ep[6];   // plastic strain at n+1
  double xi;      // hardening variable at n+1
  double sig[6];  // stress
  double c2, c3;  // tangent coefficients
  double n[6];    // unit normal of the trial deviatoric stress (0 if none)
};

// Upper triangle (by rows) of the 6x6 tangent from its coefficients, in the
// summation order of J2Plasticity.cpp:377-384.
__host__ __device__ __forceinline__ void j2_tangent_upper(const J2Param& p, double c2, double c3,
                                                          const double n[6], double D[21]) {
  const double k_twoG = 2.0 * p.G;
  int idx = 0;
#pragma unroll
  for (int a = 0; a < 6; a++) {
#pragma unroll
    for (int b = a; b < 6; b++) {
      const bool nn = (a < 3) && (b < 3);
      const bool ss = (a >= 3) && (b >= 3);
      const double IbunI = nn ? 1.0 : 0.0;
      const double IIdev = nn ? ((a == b) ? 2. / 3. : -1. / 3.) : (ss ? ((a == b) ? 0.5 : 0.0) : 0.0);
      const double NbunN = n[a] * n[b];
      double t = p.K * IbunI;
      t += k_twoG * IIdev;
      t += c2 * NbunN;
      t += c3 * (IIdev - NbunN);
      D[idx++] = t;
    }
  }
}

// Unit normal of the deviatoric part of a stress (tensor components; the
// norm counts the off-diagonals twice), 0 for a hydrostatic stress.
__host__ __device__ __forceinline__ void j2_stress_normal(const double sig[6], double n[6]) {
  const double m = (sig[0] + sig[1] + sig[2]) * (1.0 / 3.0);
  const double s[6] = {sig[0] - m, sig[1] - m, sig[2] - m, sig[3], sig[4], sig[5]};
  const double nt = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  const double inv = nt > 0.0 ? 1.0 / sqrt(nt) : 0.0;
#pragma unroll
  for (int k = 0; k < 6; k++) n[k] = s[k] * inv;
}

// The tangent re-formed from a point's stress and coefficients.
__host__ __device__ __forceinline__ void j2_tangent_from_stress(const J2Param& p, const double sig[6], double c2,
                                                                double c3, double D[21]) {
  double n[6];
  j2_stress_normal(sig, n);
  j2_tangent_upper(p, c2, c3, n, D);
}

__device__ __forceinline__ double j2_q(const J2Param& p, double xi) {
  return p.H * xi + p.sigInf + (p.sig0 - p.sigInf) * exp(-p.delta * xi);
}
__device__ __forceinline__ double j2_qprime(const J2Param& p, double xi) {
  return (p.sig0 - p.sigInf) * (-p.delta) * exp(-p.delta * xi) + p.H;
}

// eps: tensor strain (shear components already halved, J2ThreeDimensional::
// setTrialStrain :122-141).  Returns false when the local Newton iteration
// for the plastic multiplier exceeds 25 iterations (J2Plasticity.cpp:309-313);
// `o` is then left untouched, as the reference leaves its trial state.
__device__ __forceinline__ bool j2_return_map(const J2Param& p, double dt, const double eps[6],
                                              const double epn[6], double xin, J2Out& o) {
  const double root23 = sqrt(2.0 / 3.0);
  const double twoG = 2.0 * p.G;
  const double trace = eps[0] + eps[1] + eps[2];
  const double third_tr = 1. / 3. * trace;
  double dev[6];
  dev[0] = eps[0] - third_tr;
  dev[1] = eps[1] - third_tr;
  dev[2] = eps[2] - third_tr;
  dev[3] = eps[3];
  dev[4] = eps[4];
  dev[5] = eps[5];
  double ds[6];
#pragma unroll
  for (int k = 0; k < 6; k++) ds[k] = (dev[k] - epn[k]) * twoG;
  // i-major sum over the 9 tensor entries (off-diagonals counted twice)
  double nt = ds[0] * ds[0];
  nt += ds[3] * ds[3];
  nt += ds[5] * ds[5];
  nt += ds[3] * ds[3];
  nt += ds[1] * ds[1];
  nt += ds[4] * ds[4];
  nt += ds[5] * ds[5];
  nt += ds[4] * ds[4];
  nt += ds[2] * ds[2];
  const double norm_tau = sqrt(nt);
  double inv_norm_tau = 0.0;
  double n[6];
  if (norm_tau > p.tol) {
    inv_norm_tau = 1.0 / norm_tau;
#pragma unroll
    for (int k = 0; k < 6; k++) n[k] = inv_norm_tau * ds[k];
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) n[k] = 0.0;
  }
  const double phi = norm_tau - root23 * j2_q(p, xin);
  const bool visc = (p.eta > 0.0 && dt > 0.0);
  double gamma = 0.0, theta_inv = 0.0;
  if (phi > 0.0) {
    double resid = 1.0;
    int it = 0;
    while (fabs(resid) > p.tol) {
      resid = norm_tau - twoG * gamma - root23 * j2_q(p, xin + root23 * gamma);
      if (visc) resid -= (p.eta / dt) * gamma;
      double tang = -twoG - 2. / 3. * j2_qprime(p, xin + root23 * gamma);
      if (visc) tang -= (p.eta / dt);
      gamma -= (resid / tang);
      it++;
      if (it > 25) return false;
    }
    gamma *= 1.0 - 1e-08;  // J2Plasticity.cpp:318
#pragma unroll
    for (int k = 0; k < 6; k++) o.ep[k] = epn[k] + gamma * n[k];
    o.xi = xin + root23 * gamma;
#pragma unroll
    for (int k = 0; k < 6; k++) ds[k] = twoG * (dev[k] - o.ep[k]);
    double theta = twoG + 2. / 3. * j2_qprime(p, o.xi);
    if (visc) theta += (p.eta / dt);
    theta_inv = 1.0 / theta;
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) o.ep[k] = epn[k];
    o.xi = xin;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) o.sig[k] = ds[k];
  o.sig[0] += p.K * trace;
  o.sig[1] += p.K * trace;
  o.sig[2] += p.K * trace;

  const double c1 = -4.0 * p.G * p.G;
  o.c2 = c1 * theta_inv;
  o.c3 = c1 * gamma * inv_norm_tau;
#pragma unroll
  for (int k = 0; k < 6; k++) o.n[k] = n[k];
  return true;
}

}  // namespace xg
