// Oracle: CPU restatement of the reference material models on the
// state-determination path.  TEST INFRASTRUCTURE ONLY -- used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline as the checker, never
// by the product library.
//
// Each routine follows the reference arithmetic and loop order, cited by
// file:line relative to the reference's SRC/ tree.
#pragma once
#include <cmath>
#include <cfloat>
#include <cstddef>
#include <cstring>

namespace orc {

// ---------------------------------------------------------------------------
// Rank-4 identities, SRC/matrix/identity.h:1-80
// ---------------------------------------------------------------------------
inline double IbunI(int i, int j, int k, int l) { return (i == j && k == l) ? 1.0 : 0.0; }
inline double IIdev(int i, int j, int k, int l) {
  // 0.5*(d_ik d_jl + d_il d_jk) - 1/3 d_ij d_kl, tabulated in identity.h
  double sym = 0.5 * (((i == k) && (j == l)) ? 1.0 : 0.0) + 0.5 * (((i == l) && (j == k)) ? 1.0 : 0.0);
  if (i == j && k == l) {
    if (i == k) return 2. / 3.;
    return -1. / 3.;
  }
  return sym;
}

// index_map, J2Plasticity.cpp:441-488
inline void j2_index_map(int m, int& i, int& j) {
  static const int I[6] = {0, 1, 2, 0, 1, 2};
  static const int J[6] = {0, 1, 2, 1, 2, 0};
  i = I[m];
  j = J[m];
}

struct J2Params {
  double K, G, sig0, sigInf, delta, H, eta;
};

// J2Plasticity material-point object (J2Plasticity.h members used on the path)
struct J2Point {
  double epsp_n[3][3], epsp_np1[3][3];
  double xi_n, xi_np1;
  double strain[3][3], stress[3][3];
  double tangent[3][3][3][3];
  // J2Plasticity::zero, J2Plasticity.cpp:213-224: plastic state, stress and
  // strain; the tangent of the last plastic_integrator call stays (the
  // constructor and revertToStart both leave it to plastic_integrator)
  void zero() { std::memset(this, 0, offsetof(J2Point, tangent)); }
};

// q(xi), qprime(xi): J2Plasticity.cpp:425-437
inline double j2_q(const J2Params& p, double xi) {
  return p.H * xi + p.sigInf + (p.sig0 - p.sigInf) * std::exp(-p.delta * xi);
}
inline double j2_qprime(const J2Params& p, double xi) {
  return (p.sig0 - p.sigInf) * (-p.delta) * std::exp(-p.delta * xi) + p.H;
}

// J2Plasticity::plastic_integrator, J2Plasticity.cpp:228-394.
// Returns 0, or -1 when the return map exceeds 25 iterations (:309-313).
inline int j2_plastic_integrator(const J2Params& p, double dt, J2Point& s) {
  const double root23 = std::sqrt(2.0 / 3.0);
  const double tolerance = 1.0e-10 * p.sig0;  // :230
  const int max_iterations = 25;             // :243
  double dev_strain[3][3], dev_stress[3][3], normal[3][3];
  double inv_norm_tau = 0.0, tang = 0.0;

  double trace = s.strain[0][0] + s.strain[1][1] + s.strain[2][2];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) dev_strain[i][j] = s.strain[i][j];
  for (int i = 0; i < 3; i++) dev_strain[i][i] -= 1. / 3. * trace;

  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      dev_stress[i][j] = (dev_strain[i][j] - s.epsp_n[i][j]) * (2.0 * p.G);

  double norm_tau = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) norm_tau += dev_stress[i][j] * dev_stress[i][j];
  norm_tau = std::sqrt(norm_tau);

  if (norm_tau > tolerance) {
    inv_norm_tau = 1.0 / norm_tau;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) normal[i][j] = inv_norm_tau * dev_stress[i][j];
  } else {
    std::memset(normal, 0, sizeof(normal));
    inv_norm_tau = 0.0;
  }

  double phi = norm_tau - root23 * j2_q(p, s.xi_n);
  double theta_inv = 0.0, gamma = 0.0;
  if (phi > 0.0) {
    double resid = 1.0;
    int iteration_counter = 0;
    while (std::fabs(resid) > tolerance) {
      resid = norm_tau - (2.0 * p.G) * gamma - root23 * j2_q(p, s.xi_n + root23 * gamma);
      if (p.eta > 0.0 && dt > 0.0) resid -= (p.eta / dt) * gamma;
      tang = -(2.0 * p.G) - 2. / 3. * j2_qprime(p, s.xi_n + root23 * gamma);
      if (p.eta > 0.0 && dt > 0.0) tang -= (p.eta / dt);
      gamma -= (resid / tang);
      iteration_counter++;
      if (iteration_counter > max_iterations) return -1;
    }
    gamma *= 1.0 - 1e-08;  // :318
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) s.epsp_np1[i][j] = s.epsp_n[i][j] + gamma * normal[i][j];
    s.xi_np1 = s.xi_n + root23 * gamma;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        dev_stress[i][j] = (2.0 * p.G) * (dev_strain[i][j] - s.epsp_np1[i][j]);
    double theta = (2.0 * p.G) + 2. / 3. * j2_qprime(p, s.xi_np1);
    if (p.eta > 0.0 && dt > 0.0) theta += (p.eta / dt);
    theta_inv = 1.0 / theta;
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) s.epsp_np1[i][j] = s.epsp_n[i][j];
    s.xi_np1 = s.xi_n;
    gamma = 0.0;
    theta_inv = 0.0;
  }

  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) s.stress[i][j] = dev_stress[i][j];
  for (int i = 0; i < 3; i++) s.stress[i][i] += p.K * trace;

  double c1 = -4.0 * p.G * p.G;
  double c2 = c1 * theta_inv;
  double c3 = c1 * gamma * inv_norm_tau;
  for (int ii = 0; ii < 6; ii++) {
    for (int jj = 0; jj < 6; jj++) {
      int i, j, k, l;
      j2_index_map(ii, i, j);
      j2_index_map(jj, k, l);
      double NbunN = normal[i][j] * normal[k][l];
      double t = p.K * IbunI(i, j, k, l);
      t += (2.0 * p.G) * IIdev(i, j, k, l);
      t += c2 * NbunN;
      t += c3 * (IIdev(i, j, k, l) - NbunN);
      s.tangent[i][j][k][l] = t;
      s.tangent[j][i][k][l] = t;
      s.tangent[i][j][l][k] = t;
      s.tangent[j][i][l][k] = t;
    }
  }
  return 0;
}

// J2ThreeDimensional::setTrialStrain, J2ThreeDimensional.cpp:122-141
inline int j2_3d_set_trial_strain(const J2Params& p, double dt, J2Point& s, const double e[6]) {
  std::memset(s.strain, 0, sizeof(s.strain));
  s.strain[0][0] = e[0];
  s.strain[1][1] = e[1];
  s.strain[2][2] = e[2];
  s.strain[0][1] = 0.50 * e[3];
  s.strain[1][0] = s.strain[0][1];
  s.strain[1][2] = 0.50 * e[4];
  s.strain[2][1] = s.strain[1][2];
  s.strain[2][0] = 0.50 * e[5];
  s.strain[0][2] = s.strain[2][0];
  return j2_plastic_integrator(p, dt, s);
}
// J2ThreeDimensional::getStress, :189-203
inline void j2_3d_get_stress(const J2Point& s, double out[6]) {
  out[0] = s.stress[0][0];
  out[1] = s.stress[1][1];
  out[2] = s.stress[2][2];
  out[3] = s.stress[0][1];
  out[4] = s.stress[1][2];
  out[5] = s.stress[2][0];
}
// J2ThreeDimensional::getTangent, :205-234  (row-major 6x6)
inline void j2_3d_get_tangent(const J2Point& s, double D[36]) {
  for (int ii = 0; ii < 6; ii++)
    for (int jj = 0; jj < 6; jj++) {
      int i, j, k, l;
      j2_index_map(ii, i, j);
      j2_index_map(jj, k, l);
      D[ii * 6 + jj] = s.tangent[i][j][k][l];
    }
}

// J2PlaneStrain::setTrialStrain, J2PlaneStrain.cpp:124-136 (return map
// failure is ignored there: the call always returns 0)
inline int j2_ps_set_trial_strain(const J2Params& p, double dt, J2Point& s, const double e[3]) {
  std::memset(s.strain, 0, sizeof(s.strain));
  s.strain[0][0] = e[0];
  s.strain[1][1] = e[1];
  s.strain[0][1] = 0.50 * e[2];
  s.strain[1][0] = s.strain[0][1];
  j2_plastic_integrator(p, dt, s);
  return 0;
}
// J2PlaneStrain::getStress, :171-178
inline void j2_ps_get_stress(const J2Point& s, double out[3]) {
  out[0] = s.stress[0][0];
  out[1] = s.stress[1][1];
  out[2] = s.stress[0][1];
}
// J2PlaneStrain::getTangent, :180-204 (row-major 3x3)
inline void j2_ps_get_tangent(const J2Point& s, double D[9]) {
  D[0] = s.tangent[0][0][0][0];
  D[4] = s.tangent[1][1][1][1];
  D[8] = s.tangent[0][1][0][1];
  D[1] = s.tangent[0][0][1][1];
  D[3] = s.tangent[1][1][0][0];
  D[2] = s.tangent[0][0][0][1];
  D[6] = s.tangent[0][1][0][0];
  D[5] = s.tangent[1][1][0][1];
  D[7] = s.tangent[0][1][1][1];
}
// commitState, J2Plasticity.cpp:516-522; revertToLastCommit is a no-op (:524)
inline void j2_commit(J2Point& s) {
  std::memcpy(s.epsp_n, s.epsp_np1, sizeof(s.epsp_n));
  s.xi_n = s.xi_np1;
}

// ---------------------------------------------------------------------------
// Steel02 (Giuffre-Menegotto-Pinto), SRC/material/uniaxial/steel/Steel02.cpp
// ---------------------------------------------------------------------------
struct Steel02Params {
  double Fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sigini;
};
struct Steel02State {
  // committed
  double epsminP, epsmaxP, epsplP, epss0P, sigs0P, epssrP, sigsrP;
  int konP;
  double epsP, sigP, eP;
  // trial
  double epsmin, epsmax, epspl, epss0, sigs0, epsr, sigr;
  int kon;
  double eps, sig, e;
};

// Steel02::revertToStart, Steel02.cpp:69-95
inline void steel02_revert_to_start(const Steel02Params& p, Steel02State& s) {
  s.eP = p.E0;
  s.epsP = 0.0;
  s.sigP = 0.0;
  s.sig = 0.0;
  s.eps = 0.0;
  s.e = p.E0;
  s.konP = 0;
  s.epsmaxP = p.Fy / p.E0;
  s.epsminP = -s.epsmaxP;
  s.epsplP = 0.0;
  s.epss0P = 0.0;
  s.sigs0P = 0.0;
  s.epssrP = 0.0;
  s.sigsrP = 0.0;
  if (p.sigini != 0.0) {
    s.epsP = p.sigini / p.E0;
    s.sigP = p.sigini;
  }
  // trial history values are not initialised by the reference until the
  // first setTrialStrain copies them from the committed ones
  s.epsmin = s.epsminP;
  s.epsmax = s.epsmaxP;
  s.epspl = s.epsplP;
  s.epss0 = s.epss0P;
  s.sigs0 = s.sigs0P;
  s.epsr = s.epssrP;
  s.sigr = s.sigsrP;
  s.kon = s.konP;
}

// Steel02::setTrialStrain, Steel02.cpp:112-225
inline int steel02_set_trial_strain(const Steel02Params& p, Steel02State& s, double trialStrain) {
  double Esh = p.b * p.E0;
  double epsy = p.Fy / p.E0;
  if (p.sigini != 0.0) {
    double epsini = p.sigini / p.E0;
    s.eps = trialStrain + epsini;
  } else
    s.eps = trialStrain;
  double deps = s.eps - s.epsP;
  s.epsmax = s.epsmaxP;
  s.epsmin = s.epsminP;
  s.epspl = s.epsplP;
  s.epss0 = s.epss0P;
  s.sigs0 = s.sigs0P;
  s.epsr = s.epssrP;
  s.sigr = s.sigsrP;
  s.kon = s.konP;

  if (s.kon == 0 || s.kon == 3) {
    if (std::fabs(deps) < 10.0 * DBL_EPSILON) {  // :140-146
      s.e = p.E0;
      s.sig = p.sigini;
      s.kon = 3;
      return 0;
    } else {
      s.epsmax = epsy;
      s.epsmin = -epsy;
      if (deps < 0.0) {
        s.kon = 2;
        s.epss0 = s.epsmin;
        s.sigs0 = -p.Fy;
        s.epspl = s.epsmin;
      } else {
        s.kon = 1;
        s.epss0 = s.epsmax;
        s.sigs0 = p.Fy;
        s.epspl = s.epsmax;
      }
    }
  }

  if (s.kon == 2 && deps > 0.0) {  // :170-186
    s.kon = 1;
    s.epsr = s.epsP;
    s.sigr = s.sigP;
    if (s.epsP < s.epsmin) s.epsmin = s.epsP;
    double d1 = (s.epsmax - s.epsmin) / (2.0 * (p.a4 * epsy));
    double shft = 1.0 + p.a3 * std::pow(d1, 0.8);
    s.epss0 = (p.Fy * shft - Esh * epsy * shft - s.sigr + p.E0 * s.epsr) / (p.E0 - Esh);
    s.sigs0 = p.Fy * shft + Esh * (s.epss0 - epsy * shft);
    s.epspl = s.epsmax;
  } else if (s.kon == 1 && deps < 0.0) {  // :188-208
    s.kon = 2;
    s.epsr = s.epsP;
    s.sigr = s.sigP;
    if (s.epsP > s.epsmax) s.epsmax = s.epsP;
    double d1 = (s.epsmax - s.epsmin) / (2.0 * (p.a2 * epsy));
    double shft = 1.0 + p.a1 * std::pow(d1, 0.8);
    s.epss0 = (-p.Fy * shft + Esh * epsy * shft - s.sigr + p.E0 * s.epsr) / (p.E0 - Esh);
    s.sigs0 = -p.Fy * shft + Esh * (s.epss0 + epsy * shft);
    s.epspl = s.epsmin;
  }

  // :211-222
  double xi = std::fabs((s.epspl - s.epss0) / epsy);
  double R = p.R0 * (1.0 - (p.cR1 * xi) / (p.cR2 + xi));
  double epsrat = (s.eps - s.epsr) / (s.epss0 - s.epsr);
  double dum1 = 1.0 + std::pow(std::fabs(epsrat), R);
  double dum2 = std::pow(dum1, (1 / R));
  s.sig = p.b * epsrat + (1.0 - p.b) * epsrat / dum2;
  s.sig = s.sig * (s.sigs0 - s.sigr) + s.sigr;
  s.e = p.b + (1.0 - p.b) / (dum1 * dum2);
  s.e = s.e * (s.sigs0 - s.sigr) / (s.epss0 - s.epsr);
  return 0;
}
// Steel02::commitState, :247-267 (energy bookkeeping omitted: not on the path)
inline void steel02_commit(Steel02State& s) {
  s.epsminP = s.epsmin;
  s.epsmaxP = s.epsmax;
  s.epsplP = s.epspl;
  s.epss0P = s.epss0;
  s.sigs0P = s.sigs0;
  s.epssrP = s.epsr;
  s.sigsrP = s.sigr;
  s.konP = s.kon;
  s.eP = s.e;
  s.sigP = s.sig;
  s.epsP = s.eps;
}
// Steel02::revertToLastCommit, :269-285
inline void steel02_revert(Steel02State& s) {
  s.epsmin = s.epsminP;
  s.epsmax = s.epsmaxP;
  s.epspl = s.epsplP;
  s.epss0 = s.epss0P;
  s.sigs0 = s.sigs0P;
  s.epsr = s.epssrP;
  s.sigr = s.sigsrP;
  s.kon = s.konP;
  s.e = s.eP;
  s.sig = s.sigP;
  s.eps = s.epsP;
}

// ---------------------------------------------------------------------------
// Concrete02, SRC/material/uniaxial/concrete/Concrete02.cpp
// ---------------------------------------------------------------------------
struct Concrete02Params {
  double fc, epsc0, fcu, epscu, rat, ft, Ets;
};
// constructor sign normalisation, Concrete02.cpp:93-112
inline Concrete02Params concrete02_make(double fc, double epsc0, double fcu, double epscu,
                                        double rat, double ft, double Ets) {
  Concrete02Params p{fc, epsc0, fcu, epscu, rat, ft, Ets};
  if (p.fc > 0) p.fc = -p.fc;
  if (p.epsc0 > 0) p.epsc0 = -p.epsc0;
  if (p.fcu > 0) p.fcu = -p.fcu;
  if (p.epscu > 0) p.epscu = -p.epscu;
  return p;
}
struct Concrete02State {
  double ecminP, deptP, epsP, sigP, eP;
  double ecmin, dept, eps, sig, e;
};
// revertToStart, :321-337
inline void concrete02_revert_to_start(const Concrete02Params& p, Concrete02State& s) {
  s.ecminP = 0.0;
  s.deptP = 0.0;
  s.eP = 2.0 * p.fc / p.epsc0;
  s.epsP = 0.0;
  s.sigP = 0.0;
  s.eps = 0.0;
  s.sig = 0.0;
  s.e = 2.0 * p.fc / p.epsc0;
  s.ecmin = 0.0;
  s.dept = 0.0;
}
// Tens_Envlp, :420-454
inline void concrete02_tens_envlp(const Concrete02Params& p, double epsc, double& sigc, double& Ect) {
  double Ec0 = 2.0 * p.fc / p.epsc0;
  double eps0 = p.ft / Ec0;
  double epsu = p.ft * (1.0 / p.Ets + 1.0 / Ec0);
  if (epsc <= eps0) {
    sigc = epsc * Ec0;
    Ect = Ec0;
  } else {
    if (epsc <= epsu) {
      Ect = -p.Ets;
      sigc = p.ft - p.Ets * (epsc - eps0);
    } else {
      Ect = 1.0e-10;
      sigc = 0.0;
    }
  }
}
// Compr_Envlp, :457-497
inline void concrete02_compr_envlp(const Concrete02Params& p, double epsc, double& sigc, double& Ect) {
  double Ec0 = 2.0 * p.fc / p.epsc0;
  double ratLocal = epsc / p.epsc0;
  if (epsc >= p.epsc0) {
    sigc = p.fc * ratLocal * (2.0 - ratLocal);
    Ect = Ec0 * (1.0 - ratLocal);
  } else {
    if (epsc > p.epscu) {
      sigc = (p.fcu - p.fc) * (epsc - p.epsc0) / (p.epscu - p.epsc0) + p.fc;
      Ect = (p.fcu - p.fc) / (p.epscu - p.epsc0);
    } else {
      sigc = p.fcu;
      Ect = 1.0e-10;
    }
  }
}
// setTrialStrain, :166-269.  The early return at :182-183 leaves the previous
// *trial* sig/e in place.
inline int concrete02_set_trial_strain(const Concrete02Params& p, Concrete02State& s, double trialStrain) {
  double ec0 = p.fc * 2. / p.epsc0;
  s.ecmin = s.ecminP;
  s.dept = s.deptP;
  s.eps = trialStrain;
  double deps = s.eps - s.epsP;
  if (std::fabs(deps) < DBL_EPSILON) return 0;
  if (s.eps < s.ecmin) {
    concrete02_compr_envlp(p, s.eps, s.sig, s.e);
    s.ecmin = s.eps;
  } else {
    double epsr = (p.fcu - p.rat * ec0 * p.epscu) / (ec0 * (1.0 - p.rat));
    double sigmr = ec0 * epsr;
    double sigmm, dumy;
    concrete02_compr_envlp(p, s.ecmin, sigmm, dumy);
    double er = (sigmm - sigmr) / (s.ecmin - epsr);
    double ept = s.ecmin - sigmm / er;
    if (s.eps <= ept) {
      double sigmin = sigmm + er * (s.eps - s.ecmin);
      double sigmax = er * .5 * (s.eps - ept);  // `.5f` is exactly 0.5 (:221)
      s.sig = s.sigP + ec0 * deps;
      s.e = ec0;
      if (s.sig <= sigmin) {
        s.sig = sigmin;
        s.e = er;
      }
      if (s.sig >= sigmax) {
        s.sig = sigmax;
        s.e = 0.5 * er;
      }
    } else {
      double epn = ept + s.dept;
      double sicn;
      if (s.eps <= epn) {
        concrete02_tens_envlp(p, s.dept, sicn, s.e);
        if (s.dept != 0.0) {
          s.e = sicn / s.dept;
        } else {
          s.e = ec0;
        }
        s.sig = s.e * (s.eps - ept);
      } else {
        double epstmp = s.eps - ept;
        concrete02_tens_envlp(p, epstmp, s.sig, s.e);
        s.dept = s.eps - ept;
      }
    }
  }
  return 0;
}
// commitState :291-304, revertToLastCommit :306-319
inline void concrete02_commit(Concrete02State& s) {
  s.ecminP = s.ecmin;
  s.deptP = s.dept;
  s.eP = s.e;
  s.sigP = s.sig;
  s.epsP = s.eps;
}
inline void concrete02_revert(Concrete02State& s) {
  s.ecmin = s.ecminP;
  s.dept = s.deptP;
  s.e = s.eP;
  s.sig = s.sigP;
  s.eps = s.epsP;
}

}  // namespace orc
