// Uniaxial fiber materials on gfx950: Steel02 (Giuffre-Menegotto-Pinto with
// isotropic shift, Steel02.cpp:112-225) and Concrete02 (Kent-Park / Scott
// envelope, linear unload-reload and tension softening, Concrete02.cpp:166-269,
// :420-497).  State lives in structure-of-arrays pools; these functions work
// on one fiber's registers.
#pragma once
#include <cfloat>

#include "xg_common.h"

namespace xg {

// d1^0.8 of Steel02's isotropic shift, kept out of line: inlined on the
// rarely taken load-reversal path, the power's temporaries set the register
// peak of the fiber kernels (202 instead of 167 VGPRs for the 3D section
// update: two waves per SIMD instead of three).
static __device__ __attribute__((noinline)) double steel02_shift_pow(double d1) { return pow(d1, 0.8); }

struct Steel02Param {
  double Fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sigini;
};
struct Concrete02Param {
  double fc, epsc0, fcu, epscu, rat, ft, Ets;
};

// Steel02 history: epsmin, epsmax, epspl, epss0, sigs0, epsr, sigr, eps, sig, e
// (doubles, in this order) plus kon.
constexpr int kSteelHist = 10;
struct Steel02State {
  double epsmin, epsmax, epspl, epss0, sigs0, epsr, sigr, eps, sig, e;
  int kon;
};

// Steel02::setTrialStrain in two parts.  steel02_state: the history and
// branch logic (Steel02.cpp:112-190); false for the elastic return of an
// unloaded fiber (kon 3, trial stress and tangent set), true when the
// Menegotto-Pinto curve is to be evaluated by steel02_stress
// (Steel02.cpp:192-207).  Split so that a caller can store the history
// before the curve's transcendental functions run.
__device__ __forceinline__ bool steel02_state(const Steel02Param& p, const Steel02State& c, double trialStrain,
                                              Steel02State& t) {
  const double Esh = p.b * p.E0;
  const double epsy = p.Fy / p.E0;
  if (p.sigini != 0.0) {
    const double epsini = p.sigini / p.E0;
    t.eps = trialStrain + epsini;
  } else
    t.eps = trialStrain;
  const double deps = t.eps - c.eps;
  t.epsmax = c.epsmax;
  t.epsmin = c.epsmin;
  t.epspl = c.epspl;
  t.epss0 = c.epss0;
  t.sigs0 = c.sigs0;
  t.epsr = c.epsr;
  t.sigr = c.sigr;
  t.kon = c.kon;
  if (t.kon == 0 || t.kon == 3) {
    if (fabs(deps) < 10.0 * DBL_EPSILON) {  // Steel02.cpp:140-146
      t.e = p.E0;
      t.sig = p.sigini;
      t.kon = 3;
      return false;
    }
    t.epsmax = epsy;
    t.epsmin = -epsy;
    if (deps < 0.0) {
      t.kon = 2;
      t.epss0 = t.epsmin;
      t.sigs0 = -p.Fy;
      t.epspl = t.epsmin;
    } else {
      t.kon = 1;
      t.epss0 = t.epsmax;
      t.sigs0 = p.Fy;
      t.epspl = t.epsmax;
    }
  }
  // load reversal: onto the positive branch from the negative one
  // (Steel02.cpp:158-173) or the other way (:175-190); both shift the
  // asymptote by 1 + a * d1^0.8, taken by one pow for either case
  const bool up = t.kon == 2 && deps > 0.0, down = t.kon == 1 && deps < 0.0;
  if (up || down) {
    t.epsr = c.eps;
    t.sigr = c.sig;
    double d1, a;
    if (up) {
      t.kon = 1;
      if (c.eps < t.epsmin) t.epsmin = c.eps;
      d1 = (t.epsmax - t.epsmin) / (2.0 * (p.a4 * epsy));
      a = p.a3;
    } else {
      t.kon = 2;
      if (c.eps > t.epsmax) t.epsmax = c.eps;
      d1 = (t.epsmax - t.epsmin) / (2.0 * (p.a2 * epsy));
      a = p.a1;
    }
    const double shft = 1.0 + a * steel02_shift_pow(d1);
    if (up) {
      t.epss0 = (p.Fy * shft - Esh * epsy * shft - t.sigr + p.E0 * t.epsr) / (p.E0 - Esh);
      t.sigs0 = p.Fy * shft + Esh * (t.epss0 - epsy * shft);
      t.epspl = t.epsmax;
    } else {
      t.epss0 = (-p.Fy * shft + Esh * epsy * shft - t.sigr + p.E0 * t.epsr) / (p.E0 - Esh);
      t.sigs0 = -p.Fy * shft + Esh * (t.epss0 + epsy * shft);
      t.epspl = t.epsmin;
    }
  }
  return true;
}

__device__ __forceinline__ void steel02_stress(const Steel02Param& p, Steel02State& t) {
  const double epsy = p.Fy / p.E0;
  const double xi = fabs((t.epspl - t.epss0) / epsy);
  const double R = p.R0 * (1.0 - (p.cR1 * xi) / (p.cR2 + xi));
  const double epsrat = (t.eps - t.epsr) / (t.epss0 - t.epsr);
  // |epsrat|^R and dum1^(1/R) as exp(R log x): R > 0 here (R0 (1 - cR1 xi
  // / (cR2 + xi)) with cR1 < 1), |epsrat| = 0 gives exp(-inf) = 0 as pow
  // does, and the relative error, about R |log x| ulp (1e-14), is far
  // inside the parity bar; 3-4% faster fiber updates than pow
  const double dum1 = 1.0 + exp(R * log(fabs(epsrat)));
  const double dum2 = exp(log(dum1) * (1 / R));
  double sig = p.b * epsrat + (1.0 - p.b) * epsrat / dum2;
  t.sig = sig * (t.sigs0 - t.sigr) + t.sigr;
  double e = p.b + (1.0 - p.b) / (dum1 * dum2);
  t.e = e * (t.sigs0 - t.sigr) / (t.epss0 - t.epsr);
}

// Steel02::setTrialStrain: c = committed state, t = trial state written.
__device__ __forceinline__ void steel02_trial(const Steel02Param& p, const Steel02State& c, double trialStrain,
                                              Steel02State& t) {
  if (steel02_state(p, c, trialStrain, t)) steel02_stress(p, t);
}

// Concrete02 history: ecmin, dept, eps, sig, e
constexpr int kConcHist = 5;

__device__ __forceinline__ void concrete02_tens(const Concrete02Param& p, double epsc, double& sigc, double& Ect) {
  const double Ec0 = 2.0 * p.fc / p.epsc0;
  const double eps0 = p.ft / Ec0;
  const double epsu = p.ft * (1.0 / p.Ets + 1.0 / Ec0);
  if (epsc <= eps0) {
    sigc = epsc * Ec0;
    Ect = Ec0;
  } else if (epsc <= epsu) {
    Ect = -p.Ets;
    sigc = p.ft - p.Ets * (epsc - eps0);
  } else {
    Ect = 1.0e-10;
    sigc = 0.0;
  }
}

__device__ __forceinline__ void concrete02_compr(const Concrete02Param& p, double epsc, double& sigc, double& Ect) {
  const double Ec0 = 2.0 * p.fc / p.epsc0;
  const double ratLocal = epsc / p.epsc0;
  if (epsc >= p.epsc0) {
    sigc = p.fc * ratLocal * (2.0 - ratLocal);
    Ect = Ec0 * (1.0 - ratLocal);
  } else if (epsc > p.epscu) {
    sigc = (p.fcu - p.fc) * (epsc - p.epsc0) / (p.epscu - p.epsc0) + p.fc;
    Ect = (p.fcu - p.fc) / (p.epscu - p.epsc0);
  } else {
    sigc = p.fcu;
    Ect = 1.0e-10;
  }
}

// Concrete02::setTrialStrain.  c* = committed ecmin, dept, eps, sig; the trial
// sig/e are in-out: on the early return (|deps| < DBL_EPSILON, :182-183) they
// keep their previous trial values.
__device__ __forceinline__ void concrete02_trial(const Concrete02Param& p, double cecmin, double cdept,
                                                 double ceps, double csig, double trialStrain, double& ecmin,
                                                 double& dept, double& eps, double& sig, double& e) {
  const double ec0 = p.fc * 2. / p.epsc0;
  ecmin = cecmin;
  dept = cdept;
  eps = trialStrain;
  const double deps = eps - ceps;
  if (fabs(deps) < DBL_EPSILON) return;
  if (eps < ecmin) {
    concrete02_compr(p, eps, sig, e);
    ecmin = eps;
    return;
  }
  const double epsr = (p.fcu - p.rat * ec0 * p.epscu) / (ec0 * (1.0 - p.rat));
  const double sigmr = ec0 * epsr;
  double sigmm, dumy;
  concrete02_compr(p, ecmin, sigmm, dumy);
  const double er = (sigmm - sigmr) / (ecmin - epsr);
  const double ept = ecmin - sigmm / er;
  if (eps <= ept) {
    const double sigmin = sigmm + er * (eps - ecmin);
    const double sigmax = er * .5 * (eps - ept);
    sig = csig + ec0 * deps;
    e = ec0;
    if (sig <= sigmin) {
      sig = sigmin;
      e = er;
    }
    if (sig >= sigmax) {
      sig = sigmax;
      e = 0.5 * er;
    }
  } else {
    const double epn = ept + dept;
    if (eps <= epn) {
      double sicn;
      concrete02_tens(p, dept, sicn, e);
      if (dept != 0.0) e = sicn / dept;
      else e = ec0;
      sig = e * (eps - ept);
    } else {
      const double epstmp = eps - ept;
      concrete02_tens(p, epstmp, sig, e);
      dept = eps - ept;
    }
  }
}

}  // namespace xg
