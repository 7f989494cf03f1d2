// Reverse-mode (adjoint) derivatives of the Steel02 and Concrete02 trial
// updates of xg_uniaxial.h: for one fiber, given the committed history, the
// strain and the adjoint of the trial outputs, return the adjoints of the
// committed inputs, of the strain and of the varying parameters (Steel02 Fy,
// E0, b; Concrete02 fc).  The forward values are recomputed first, so the
// derivative follows the branch the trial update takes (Steel02.cpp:112-225,
// Concrete02.cpp:166-269 and the envelopes :420-497).
//
// Histories are in pool order: Steel02 epsmin, epsmax, epspl, epss0, sigs0,
// epsr, sigr, eps, sig, e; Concrete02 ecmin, dept, eps, sig, e.  The tangent
// outputs (e) never enter a resisting force, so their adjoints are not
// propagated (their incoming adjoints are ignored).
#pragma once
#include <cfloat>
#include <cmath>

#include "xg_uniaxial.h"

namespace xg {

__host__ __device__ inline void steel02_vjp(const Steel02Param& p, const double* c, int ckon, double strain,
                                            const double* tb, double* cb, double* thb, double& sb) {
  const double Fy = p.Fy, E0 = p.E0, bb = p.b;
  const double Esh = bb * E0;
  const double epsy = Fy / E0;
  const double eps = p.sigini != 0.0 ? strain + p.sigini / E0 : strain;
  const double deps = eps - c[7];
  double epsmin = c[0], epsmax = c[1], epspl = c[2], epss0 = c[3], sigs0 = c[4], epsr = c[5], sigr = c[6];
  int kon = ckon;
  // path: 0 history carried over, 1 no strain change from the virgin state,
  // 2 first loading, 3 reversal 2->1, 4 reversal 1->2
  int path = 0;
  bool neg = false, sel = false;
  double d1 = 0, shft = 0;
  if (kon == 0 || kon == 3) {
    if (fabs(deps) < 10.0 * DBL_EPSILON) {
      path = 1;
    } else {
      path = 2;
      epsmax = epsy;
      epsmin = -epsy;
      neg = deps < 0.0;
      kon = neg ? 2 : 1;
      epss0 = neg ? epsmin : epsmax;
      sigs0 = neg ? -Fy : Fy;
      epspl = epss0;
    }
  }
  if (path != 1 && kon == 2 && deps > 0.0) {
    path = 3;
    epsr = c[7];
    sigr = c[8];
    sel = c[7] < epsmin;
    if (sel) epsmin = c[7];
    d1 = (epsmax - epsmin) / (2.0 * (p.a4 * epsy));
    shft = 1.0 + p.a3 * pow(d1, 0.8);
    epss0 = (Fy * shft - Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
    sigs0 = Fy * shft + Esh * (epss0 - epsy * shft);
    epspl = epsmax;
  } else if (path != 1 && kon == 1 && deps < 0.0) {
    path = 4;
    epsr = c[7];
    sigr = c[8];
    sel = c[7] > epsmax;
    if (sel) epsmax = c[7];
    d1 = (epsmax - epsmin) / (2.0 * (p.a2 * epsy));
    shft = 1.0 + p.a1 * pow(d1, 0.8);
    epss0 = (-Fy * shft + Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
    sigs0 = -Fy * shft + Esh * (epss0 + epsy * shft);
    epspl = epsmin;
  }
  for (int k = 0; k < 10; k++) cb[k] = 0.0;
  double b_epsmin = tb[0], b_epsmax = tb[1], b_epspl = tb[2], b_epss0 = tb[3], b_sigs0 = tb[4];
  double b_epsr = tb[5], b_sigr = tb[6], b_eps = tb[7];
  const double b_sig = tb[8];
  double b_Fy = 0, b_E0 = 0, b_b = 0, b_epsy = 0, b_Esh = 0;
  if (path == 1) {
    // trial = committed history, sig = sigini
    cb[0] = b_epsmin;
    cb[1] = b_epsmax;
    cb[2] = b_epspl;
    cb[3] = b_epss0;
    cb[4] = b_sigs0;
    cb[5] = b_epsr;
    cb[6] = b_sigr;
  } else {
    // ---- Menegotto-Pinto curve (Steel02.cpp:211-222) ----
    const double q = (epspl - epss0) / epsy;
    const double xi = fabs(q);
    const double R = p.R0 * (1.0 - (p.cR1 * xi) / (p.cR2 + xi));
    const double den = epss0 - epsr;
    const double er = (eps - epsr) / den;
    const double aer = fabs(er);
    const double pw = pow(aer, R);
    const double dum1 = 1.0 + pw;
    const double dum2 = pow(dum1, (1 / R));
    const double sstar = bb * er + (1.0 - bb) * er / dum2;
    const double b_sstar = b_sig * (sigs0 - sigr);
    b_sigs0 += b_sig * sstar;
    b_sigr += b_sig * (1.0 - sstar);
    b_b += b_sstar * (er - er / dum2);
    double b_er = b_sstar * (bb + (1.0 - bb) / dum2);
    const double b_dum2 = -b_sstar * (1.0 - bb) * er / (dum2 * dum2);
    const double b_dum1 = b_dum2 * dum2 / (R * dum1);
    double b_R = -b_dum2 * dum2 * log(dum1) / (R * R);
    if (aer > 0.0) {
      b_er += b_dum1 * R * (pw / aer) * (er < 0.0 ? -1.0 : 1.0);
      b_R += b_dum1 * pw * log(aer);
    }
    b_eps += b_er / den;
    b_epsr += b_er * (er - 1.0) / den;
    b_epss0 += -b_er * er / den;
    const double b_xi = -b_R * p.R0 * p.cR1 * p.cR2 / ((p.cR2 + xi) * (p.cR2 + xi));
    const double b_q = b_xi * (q < 0.0 ? -1.0 : 1.0);
    b_epspl += b_q / epsy;
    b_epss0 += -b_q / epsy;
    b_epsy += -b_q * q / epsy;
    // ---- history update of the branch taken ----
    if (path == 0) {
      cb[0] = b_epsmin;
      cb[1] = b_epsmax;
      cb[2] = b_epspl;
      cb[3] = b_epss0;
      cb[4] = b_sigs0;
      cb[5] = b_epsr;
      cb[6] = b_sigr;
    } else if (path == 2) {
      const double sg = neg ? -1.0 : 1.0;
      b_epsy += b_epsmax - b_epsmin + sg * (b_epss0 + b_epspl);
      b_Fy += sg * b_sigs0;
      cb[5] = b_epsr;
      cb[6] = b_sigr;
    } else {
      const double s = path == 3 ? 1.0 : -1.0;  // sign of Fy in the shifted asymptote
      const double ac = path == 3 ? p.a3 : p.a1, aw = path == 3 ? p.a4 : p.a2;
      // sigs0 = s Fy shft + Esh (epss0 - s epsy shft)
      double b_shft = b_sigs0 * (s * Fy - s * Esh * epsy);
      b_Fy += b_sigs0 * s * shft;
      b_Esh += b_sigs0 * (epss0 - s * epsy * shft);
      b_epss0 += b_sigs0 * Esh;
      b_epsy += -b_sigs0 * s * Esh * shft;
      // epss0 = (s Fy shft - s Esh epsy shft - sigr + E0 epsr) / (E0 - Esh)
      const double Dn = E0 - Esh;
      const double b_N = b_epss0 / Dn;
      const double b_Dn = -b_epss0 * epss0 / Dn;
      b_Fy += b_N * s * shft;
      b_shft += b_N * (s * Fy - s * Esh * epsy);
      b_Esh += -b_N * s * epsy * shft - b_Dn;
      b_epsy += -b_N * s * Esh * shft;
      b_sigr += -b_N;
      b_E0 += b_N * epsr + b_Dn;
      b_epsr += b_N * E0;
      // shft = 1 + ac d1^0.8, d1 = (epsmax - epsmin) / (2 aw epsy)
      const double b_d1 = b_shft * ac * 0.8 * pow(d1, -0.2);
      const double w = 2.0 * (aw * epsy);
      b_epsmax += b_d1 / w;
      b_epsmin += -b_d1 / w;
      b_epsy += -b_d1 * d1 / epsy;
      if (path == 3) {
        b_epsmax += b_epspl;  // epspl = epsmax
        if (sel) cb[7] += b_epsmin;  // epsmin = min(epsmin, eps committed)
        else cb[0] += b_epsmin;
        cb[1] += b_epsmax;
      } else {
        b_epsmin += b_epspl;  // epspl = epsmin
        if (sel) cb[7] += b_epsmax;
        else cb[1] += b_epsmax;
        cb[0] += b_epsmin;
      }
      cb[7] += b_epsr;  // epsr = committed eps
      cb[8] += b_sigr;  // sigr = committed sig
    }
  }
  sb = b_eps;
  if (p.sigini != 0.0) b_E0 += -b_eps * p.sigini / (E0 * E0);
  b_b += b_Esh * E0;
  b_E0 += b_Esh * bb;
  b_Fy += b_epsy / E0;
  b_E0 += -b_epsy * Fy / (E0 * E0);
  thb[0] = b_Fy;
  thb[1] = b_E0;
  thb[2] = b_b;
}

// Concrete02 envelopes with their partials d/dx and d/dfc
__host__ __device__ inline double conc_tens_d(const Concrete02Param& p, double x, double& dx, double& dfc) {
  const double Ec0 = 2.0 * p.fc / p.epsc0;
  const double eps0 = p.ft / Ec0;
  const double epsu = p.ft * (1.0 / p.Ets + 1.0 / Ec0);
  if (x <= eps0) {
    dx = Ec0;
    dfc = x * 2.0 / p.epsc0;
    return x * Ec0;
  }
  if (x <= epsu) {
    dx = -p.Ets;
    dfc = p.Ets * (-eps0 / Ec0) * (2.0 / p.epsc0);
    return p.ft - p.Ets * (x - eps0);
  }
  dx = 0.0;
  dfc = 0.0;
  return 0.0;
}
__host__ __device__ inline double conc_compr_d(const Concrete02Param& p, double x, double& dx, double& dfc) {
  const double r = x / p.epsc0;
  if (x >= p.epsc0) {
    dx = p.fc * (2.0 - 2.0 * r) / p.epsc0;
    dfc = r * (2.0 - r);
    return p.fc * r * (2.0 - r);
  }
  if (x > p.epscu) {
    const double w = p.epscu - p.epsc0;
    dx = (p.fcu - p.fc) / w;
    dfc = 1.0 - (x - p.epsc0) / w;
    return (p.fcu - p.fc) * (x - p.epsc0) / w + p.fc;
  }
  dx = 0.0;
  dfc = 0.0;
  return p.fcu;
}

__host__ __device__ inline void concrete02_vjp(const Concrete02Param& p, const double* c, double strain,
                                               const double* tb, double* cb, double& fcb, double& sb) {
  const double fc = p.fc;
  const double ec0 = fc * 2. / p.epsc0;
  const double ecmin = c[0], dept = c[1], eps = strain;
  const double deps = eps - c[2];
  for (int k = 0; k < 5; k++) cb[k] = 0.0;
  fcb = 0.0;
  const double b_ecmin_o = tb[0], b_dept_o = tb[1], b_eps_o = tb[2], b_sig = tb[3];
  if (fabs(deps) < DBL_EPSILON) {
    // trial keeps the previous (committed) sig / e
    cb[0] = b_ecmin_o;
    cb[1] = b_dept_o;
    cb[3] = b_sig;
    sb = b_eps_o;
    return;
  }
  double dx, dfc;
  if (eps < ecmin) {
    conc_compr_d(p, eps, dx, dfc);
    sb = b_eps_o + b_ecmin_o + b_sig * dx;
    fcb = b_sig * dfc;
    cb[1] = b_dept_o;
    return;
  }
  const double epsr = (p.fcu - p.rat * ec0 * p.epscu) / (ec0 * (1.0 - p.rat));
  const double sigmr = ec0 * epsr;
  double dmm_x, dmm_fc;
  const double sigmm = conc_compr_d(p, ecmin, dmm_x, dmm_fc);
  const double Dd = ecmin - epsr;
  const double er = (sigmm - sigmr) / Dd;
  const double ept = ecmin - sigmm / er;
  double b_ec0 = 0, b_sigmm = 0, b_er = 0, b_ept = 0, b_ecmin = b_ecmin_o, b_dept = 0, b_eps = b_eps_o;
  if (eps <= ept) {
    const double sigmin = sigmm + er * (eps - ecmin);
    const double sigmax = er * .5 * (eps - ept);
    double sig = c[3] + ec0 * deps;
    int sel = 0;
    if (sig <= sigmin) {
      sig = sigmin;
      sel = 1;
    }
    if (sig >= sigmax) sel = 2;
    if (sel == 0) {
      cb[3] += b_sig;
      b_ec0 += b_sig * deps;
      b_eps += b_sig * ec0;
      cb[2] += -b_sig * ec0;
    } else if (sel == 1) {
      b_sigmm += b_sig;
      b_er += b_sig * (eps - ecmin);
      b_eps += b_sig * er;
      b_ecmin += -b_sig * er;
    } else {
      b_er += b_sig * .5 * (eps - ept);
      b_eps += b_sig * .5 * er;
      b_ept += -b_sig * .5 * er;
    }
    b_dept += b_dept_o;
  } else if (eps <= ept + dept) {
    double e;
    double sicn = 0, tdx = 0, tdfc = 0;
    if (dept != 0.0) {
      sicn = conc_tens_d(p, dept, tdx, tdfc);
      e = sicn / dept;
    } else {
      e = ec0;
    }
    const double b_e = b_sig * (eps - ept);
    b_eps += b_sig * e;
    b_ept += -b_sig * e;
    if (dept != 0.0) {
      const double b_sicn = b_e / dept;
      b_dept += -b_e * sicn / (dept * dept) + b_sicn * tdx;
      fcb += b_sicn * tdfc;
    } else {
      b_ec0 += b_e;
    }
    b_dept += b_dept_o;
  } else {
    const double x = eps - ept;
    double tdx, tdfc;
    conc_tens_d(p, x, tdx, tdfc);
    const double b_x = b_sig * tdx + b_dept_o;  // dept = eps - ept
    fcb += b_sig * tdfc;
    b_eps += b_x;
    b_ept += -b_x;
  }
  // ept = ecmin - sigmm / er
  b_ecmin += b_ept;
  b_sigmm += -b_ept / er;
  b_er += b_ept * sigmm / (er * er);
  // er = (sigmm - sigmr) / (ecmin - epsr)
  b_sigmm += b_er / Dd;
  const double b_sigmr = -b_er / Dd;
  b_ecmin += -b_er * er / Dd;
  double b_epsr = b_er * er / Dd;
  // sigmm = compr(ecmin)
  b_ecmin += b_sigmm * dmm_x;
  fcb += b_sigmm * dmm_fc;
  // sigmr = ec0 epsr; epsr = (fcu - rat ec0 epscu) / (ec0 (1 - rat))
  b_ec0 += b_sigmr * epsr;
  b_epsr += b_sigmr * ec0;
  b_ec0 += b_epsr * (-p.fcu / (ec0 * ec0 * (1.0 - p.rat)));
  fcb += b_ec0 * 2.0 / p.epsc0;
  cb[0] += b_ecmin;
  cb[1] += b_dept;
  sb = b_eps;
}

}  // namespace xg
