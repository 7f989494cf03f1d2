// Oracle: forward-mode (tangent-linear) restatement of the Steel02 and
// Concrete02 trial-state updates, for the direct-differentiation (DDM)
// cross-check of the GPU adjoint.
//
// TEST INFRASTRUCTURE ONLY (the checker for the GPU library).
//
// The reference has no DDM for these materials: neither Steel02 nor
// Concrete02 overrides getStressSensitivity / commitSensitivity
// (UniaxialMaterial.cpp:340-379 returns 0), so the material partials are
// restated here.  The functions below are the oracle's
// steel02_set_trial_strain / concrete02_set_trial_strain (orc_materials.hpp,
// following Steel02.cpp:112-225 and Concrete02.cpp:166-269) written over a
// scalar type T: T = double reproduces them bit for bit (checked by the
// tests), T = Dual carries one directional derivative through every
// operation (forward-mode automatic differentiation), so the derivative is
// taken along exactly the branch the value takes.
#pragma once
#include <cfloat>
#include <cmath>

#include "orc_materials.hpp"

namespace orc {

struct Dual {
  double v = 0, d = 0;
  Dual() {}
  Dual(double x) : v(x), d(0) {}
  Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline Dual dabs(Dual a) { return a.v < 0 ? -a : a; }
inline double dabs(double a) { return std::fabs(a); }
// x^y, both varying (x > 0 or x = 0 with y > 1 on the paths used here)
inline Dual dpow(Dual x, Dual y) {
  const double p = std::pow(x.v, y.v);
  if (x.v == 0.0) return {p, 0.0};
  return {p, y.v * std::pow(x.v, y.v - 1.0) * x.d + p * std::log(x.v) * y.d};
}
inline double dpow(double x, double y) { return std::pow(x, y); }
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }

// Steel02 history in pool order: epsmin, epsmax, epspl, epss0, sigs0, epsr,
// sigr, eps, sig, e (+ kon)
template <class T>
struct SteelH {
  T epsmin, epsmax, epspl, epss0, sigs0, epsr, sigr, eps, sig, e;
  int kon;
};
// the varying Steel02 parameters; the others are constants of the material
template <class T>
struct SteelTheta {
  T Fy, E0, b;
};

template <class T>
inline void steel02_trial_t(const Steel02Params& p, const SteelTheta<T>& th, const SteelH<T>& c, T trialStrain,
                            SteelH<T>& t) {
  const T Esh = th.b * th.E0;
  const T epsy = th.Fy / th.E0;
  if (p.sigini != 0.0) {
    const T epsini = T(p.sigini) / th.E0;
    t.eps = trialStrain + epsini;
  } else
    t.eps = trialStrain;
  const T deps = t.eps - c.eps;
  t.epsmax = c.epsmax;
  t.epsmin = c.epsmin;
  t.epspl = c.epspl;
  t.epss0 = c.epss0;
  t.sigs0 = c.sigs0;
  t.epsr = c.epsr;
  t.sigr = c.sigr;
  t.kon = c.kon;
  if (t.kon == 0 || t.kon == 3) {
    if (std::fabs(val(deps)) < 10.0 * DBL_EPSILON) {
      t.e = th.E0;
      t.sig = T(p.sigini);
      t.kon = 3;
      return;
    }
    t.epsmax = epsy;
    t.epsmin = -epsy;
    if (val(deps) < 0.0) {
      t.kon = 2;
      t.epss0 = t.epsmin;
      t.sigs0 = -th.Fy;
      t.epspl = t.epsmin;
    } else {
      t.kon = 1;
      t.epss0 = t.epsmax;
      t.sigs0 = th.Fy;
      t.epspl = t.epsmax;
    }
  }
  if (t.kon == 2 && val(deps) > 0.0) {
    t.kon = 1;
    t.epsr = c.eps;
    t.sigr = c.sig;
    if (c.eps < t.epsmin) t.epsmin = c.eps;
    const T d1 = (t.epsmax - t.epsmin) / (T(2.0) * (T(p.a4) * epsy));
    const T shft = T(1.0) + T(p.a3) * dpow(d1, T(0.8));
    t.epss0 = (th.Fy * shft - Esh * epsy * shft - t.sigr + th.E0 * t.epsr) / (th.E0 - Esh);
    t.sigs0 = th.Fy * shft + Esh * (t.epss0 - epsy * shft);
    t.epspl = t.epsmax;
  } else if (t.kon == 1 && val(deps) < 0.0) {
    t.kon = 2;
    t.epsr = c.eps;
    t.sigr = c.sig;
    if (c.eps > t.epsmax) t.epsmax = c.eps;
    const T d1 = (t.epsmax - t.epsmin) / (T(2.0) * (T(p.a2) * epsy));
    const T shft = T(1.0) + T(p.a1) * dpow(d1, T(0.8));
    t.epss0 = (-th.Fy * shft + Esh * epsy * shft - t.sigr + th.E0 * t.epsr) / (th.E0 - Esh);
    t.sigs0 = -th.Fy * shft + Esh * (t.epss0 + epsy * shft);
    t.epspl = t.epsmin;
  }
  const T xi = dabs((t.epspl - t.epss0) / epsy);
  const T R = T(p.R0) * (T(1.0) - (T(p.cR1) * xi) / (T(p.cR2) + xi));
  const T epsrat = (t.eps - t.epsr) / (t.epss0 - t.epsr);
  const T dum1 = T(1.0) + dpow(dabs(epsrat), R);
  const T dum2 = dpow(dum1, T(1.0) / R);
  T sig = th.b * epsrat + (T(1.0) - th.b) * epsrat / dum2;
  t.sig = sig * (t.sigs0 - t.sigr) + t.sigr;
  T e = th.b + (T(1.0) - th.b) / (dum1 * dum2);
  t.e = e * (t.sigs0 - t.sigr) / (t.epss0 - t.epsr);
}

// Concrete02 history in pool order: ecmin, dept, eps, sig, e
template <class T>
struct ConcH {
  T ecmin, dept, eps, sig, e;
};

template <class T>
inline void concrete02_tens_t(const Concrete02Params& p, T fc, T epsc, T& sigc, T& Ect) {
  const T Ec0 = T(2.0) * fc / T(p.epsc0);
  const T eps0 = T(p.ft) / Ec0;
  const T epsu = T(p.ft) * (T(1.0) / T(p.Ets) + T(1.0) / Ec0);
  if (epsc <= eps0) {
    sigc = epsc * Ec0;
    Ect = Ec0;
  } else if (epsc <= epsu) {
    Ect = T(-p.Ets);
    sigc = T(p.ft) - T(p.Ets) * (epsc - eps0);
  } else {
    Ect = T(1.0e-10);
    sigc = T(0.0);
  }
}

template <class T>
inline void concrete02_compr_t(const Concrete02Params& p, T fc, T epsc, T& sigc, T& Ect) {
  const T Ec0 = T(2.0) * fc / T(p.epsc0);
  const T ratLocal = epsc / T(p.epsc0);
  if (val(epsc) >= p.epsc0) {
    sigc = fc * ratLocal * (T(2.0) - ratLocal);
    Ect = Ec0 * (T(1.0) - ratLocal);
  } else if (val(epsc) > p.epscu) {
    sigc = (T(p.fcu) - fc) * (epsc - T(p.epsc0)) / T(p.epscu - p.epsc0) + fc;
    Ect = (T(p.fcu) - fc) / T(p.epscu - p.epsc0);
  } else {
    sigc = T(p.fcu);
    Ect = T(1.0e-10);
  }
}

// c = committed (ecmin, dept, eps, sig used), prev = previous trial sig / e
// (kept on the early return), fc the varying parameter
template <class T>
inline void concrete02_trial_t(const Concrete02Params& p, T fc, const ConcH<T>& c, T prev_sig, T prev_e,
                               T trialStrain, ConcH<T>& t) {
  const T ec0 = fc * T(2.) / T(p.epsc0);
  t.ecmin = c.ecmin;
  t.dept = c.dept;
  t.eps = trialStrain;
  t.sig = prev_sig;
  t.e = prev_e;
  const T deps = t.eps - c.eps;
  if (std::fabs(val(deps)) < DBL_EPSILON) return;
  if (t.eps < t.ecmin) {
    concrete02_compr_t(p, fc, t.eps, t.sig, t.e);
    t.ecmin = t.eps;
    return;
  }
  const T epsr = (T(p.fcu) - T(p.rat) * ec0 * T(p.epscu)) / (ec0 * T(1.0 - p.rat));
  const T sigmr = ec0 * epsr;
  T sigmm, dumy;
  concrete02_compr_t(p, fc, t.ecmin, sigmm, dumy);
  const T er = (sigmm - sigmr) / (t.ecmin - epsr);
  const T ept = t.ecmin - sigmm / er;
  if (t.eps <= ept) {
    const T sigmin = sigmm + er * (t.eps - t.ecmin);
    const T sigmax = er * T(.5) * (t.eps - ept);
    t.sig = c.sig + ec0 * deps;
    t.e = ec0;
    if (t.sig <= sigmin) {
      t.sig = sigmin;
      t.e = er;
    }
    if (t.sig >= sigmax) {
      t.sig = sigmax;
      t.e = T(0.5) * er;
    }
  } else {
    const T epn = ept + t.dept;
    if (t.eps <= epn) {
      T sicn;
      concrete02_tens_t(p, fc, t.dept, sicn, t.e);
      if (val(t.dept) != 0.0) t.e = sicn / t.dept;
      else t.e = ec0;
      t.sig = t.e * (t.eps - ept);
    } else {
      const T epstmp = t.eps - ept;
      concrete02_tens_t(p, fc, epstmp, t.sig, t.e);
      t.dept = t.eps - ept;
    }
  }
}

}  // namespace orc
