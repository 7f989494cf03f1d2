// J2 plasticity material-point update for gfx950: backward-Euler radial
// return with linear + saturation isotropic hardening and optional linear
// viscosity, following J2Plasticity::plastic_integrator
// (SRC/material/plastic/J2Plasticity.cpp:228-394) in symmetric storage.
//
// Strain, plastic strain and stress are symmetric tensors held as six tensor
// components in the order 00,11,22,01,12,20 (J2ThreeDimensional.cpp:189-203);
// the consistent tangent is the Voigt 6x6 matrix of J2ThreeDimensional::
// getTangent (:205-234) -- symmetric, so only its upper triangle is formed.
#pragma once
#include "xg_common.h"

namespace xg {

struct J2Out {
  double ep[6];   // plastic strain at n+1
  double xi;      // hardening variable at n+1
  double sig[6];  // stress
  double D[21];   // upper triangle of the 6x6 tangent, by rows
};

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
  const double c2 = c1 * theta_inv;
  const double c3 = c1 * gamma * inv_norm_tau;
  const double k_twoG = twoG;
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
      o.D[idx++] = t;
    }
  }
  return true;
}

}  // namespace xg
