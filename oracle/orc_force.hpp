// CPU oracle of ForceBeamColumn2d / ForceBeamColumn3d (test infrastructure
// only): the element-level compatibility iteration of
// SRC/element/Frame/Other/Force/ForceBeamColumn2d.cpp:559-934 (3D:
// ForceBeamColumn3d.cpp:588-1040) over the oracle's fiber sections, with
// Gauss-Lobatto integration (LobattoBeamIntegration.cpp:113-293, the
// reference's own truncated constants), a linear coordinate transformation,
// no element loads and no mass.
#pragma once
#include <array>
#include <cfloat>
#include <cmath>
#include <vector>

#include "orc_frame.hpp"

namespace orc {

// Gauss-Lobatto locations on [-1, 1] and weights on [-1, 1] for 2..10
// points, as LobattoBeamIntegration::getSectionLocations / Weights tabulate
// them (the section locations become 0.5 (x + 1), the weights 0.5 w).
inline bool lobatto_rule(int n, double* xi, double* wt) {
  static const double X[11][10] = {
      {}, {},
      {-1.0, 1.0},
      {-1.0, 0.0, 1.0},
      {-1.0, -0.44721360, 0.44721360, 1.0},
      {-1.0, -0.65465367, 0.0, 0.65465367, 1.0},
      {-1.0, -0.7650553239, -0.2852315164, 0.2852315164, 0.7650553239, 1.0},
      {-1.0, -0.8302238962, -0.4688487934, 0.0, 0.4688487934, 0.8302238962, 1.0},
      {-1.0, -0.8717401485, -0.5917001814, -0.2092992179, 0.2092992179, 0.5917001814, 0.8717401485, 1.0},
      {-1.0, -0.8997579954, -0.6771862795, -0.3631174638, 0.0, 0.3631174638, 0.6771862795, 0.8997579954, 1.0},
      {-1.0, -0.9195339082, -0.7387738651, -0.4779249498, -0.1652789577, 0.1652789577, 0.4779249498,
       0.7387738651, 0.9195339082, 1.0}};
  static const double W[11][10] = {
      {}, {},
      {1.0, 1.0},
      {0.333333333333333, 1.333333333333333, 0.333333333333333},
      {0.166666666666667, 0.833333333333333, 0.833333333333333, 0.166666666666667},
      {0.1, 0.5444444444, 0.7111111111, 0.5444444444, 0.1},
      {0.06666666667, 0.3784749562, 0.5548583770, 0.5548583770, 0.3784749562, 0.06666666667},
      {0.04761904762, 0.2768260473, 0.4317453812, 0.4876190476, 0.4317453812, 0.2768260473, 0.04761904762},
      {0.03571428571, 0.2107042271, 0.3411226924, 0.4124587946, 0.4124587946, 0.3411226924, 0.2107042271,
       0.03571428571},
      {0.02777777778, 0.1654953615, 0.2745387125, 0.3464285109, 0.3715192743, 0.3464285109, 0.2745387125,
       0.1654953615, 0.02777777778},
      {0.02222222222, 0.1333059908, 0.2248893421, 0.2920426836, 0.3275397611, 0.3275397611, 0.2920426836,
       0.2248893421, 0.1333059908, 0.02222222222}};
  if (n < 2 || n > 10) return false;
  for (int i = 0; i < n; i++) {
    xi[i] = 0.5 * (X[n][i] + 1.0);
    wt[i] = W[n][i] * 0.5;
  }
  return true;
}

// Inverse of a small dense matrix (row-major) by Gauss-Jordan elimination
// with partial pivoting; -1 for a singular matrix.
inline int small_inverse(int n, const double* A, double* Ainv) {
  double a[6][12];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) a[i][j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int i = c + 1; i < n; i++)
      if (std::fabs(a[i][c]) > std::fabs(a[piv][c])) piv = i;
    if (a[piv][c] == 0.0) return -1;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) std::swap(a[c][j], a[piv][j]);
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * n; j++) a[c][j] *= inv;
    for (int i = 0; i < n; i++)
      if (i != c) {
        const double f = a[i][c];
        for (int j = 0; j < 2 * n; j++) a[i][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Ainv[i * n + j] = a[i][n + j];
  return 0;
}

struct ForceElement {
  FrameElement fe;  // geometry, transformation and the fiber sections
  int maxIters = 10;
  double tol = 1e-12;
  double xi[10] = {0}, wt[10] = {0};
  int initialFlag = 0;
  double Se[6] = {0}, kv[36] = {0}, Secommit[6] = {0}, kvcommit[36] = {0};
  // basic displacement at the previous update (the element's view of the
  // nodes' incrDeltaDisp) and at the last commit
  double vprev[6] = {0}, vprev_c[6] = {0};
  std::vector<std::array<double, 4>> vs, vscommit, Ssr;
  std::vector<std::array<double, 16>> fs;

  int NB() const { return fe.is3d ? 6 : 3; }
  int NS() const { return fe.is3d ? 4 : 2; }

  int init(bool three_d, const double ci[3], const double cj[3], const double* vecxz, const FiberSectionDef& sd,
           int nip, int max_iter, double tol_) {
    if (three_d && sd.GJ < 0) return -4;  // the flexibility of a section without torsion is singular
    int r = fe.init(three_d, ci, cj, vecxz, sd, nip, 0.0);
    if (r < 0) return r;
    if (!lobatto_rule(nip, xi, wt)) return -5;
    maxIters = max_iter;
    tol = tol_;
    std::array<double, 4> z4{};
    std::array<double, 16> z16{};
    vs.assign(nip, z4);
    vscommit.assign(nip, z4);
    Ssr.assign(nip, z4);
    fs.assign(nip, z16);
    // Domain::addElement -> setDomain + update() at zero displacement
    double ug[12] = {0};
    return update(ug) < 0 ? -6 : 0;
  }

  // section tangent (NS x NS, row-major) of section i
  void section_ks(int i, double* k) const {
    const int ns = NS();
    for (int a = 0; a < ns; a++)
      for (int b = 0; b < ns; b++) k[a * ns + b] = fe.secs[i].ks(a, b);
  }
  // SectionForceDeformation::getInitialFlexibility: inverse of the initial
  // tangent (fiber initial tangents, FiberSection2d.cpp:271-295; 3D alike)
  void initial_flexibility(int i, double* f0) const {
    const FiberSection& s = fe.secs[i];
    double k[16] = {0};
    for (size_t q = 0; q < s.def->fibers.size(); q++) {
      const FiberDef& f = s.def->fibers[q];
      const double E = f.kind == 1 ? f.steel.E0 : (f.kind == 2 ? 2.0 * f.conc.fc / f.conc.epsc0 : f.E);
      const double y = f.y - s.yBar, A = f.A;
      if (!fe.is3d) {
        const double ks0 = E * A, ks1 = ks0 * -y;
        k[0] += ks0;
        k[1] += ks1;
        k[3] += ks1 * -y;
      } else {
        const double z = f.z - s.zBar, EA = E * A;
        k[0] += EA;
        k[1] += -y * EA;
        k[2] += z * EA;
        k[5] += y * y * EA;
        k[6] += -y * z * EA;
        k[10] += z * z * EA;
      }
    }
    if (!fe.is3d) {
      k[2] = k[1];
    } else {
      k[4] = k[1];
      k[8] = k[2];
      k[9] = k[6];
      k[15] = s.def->GJ;
    }
    small_inverse(NS(), k, f0);
  }

  // ForceBeamColumn2d::update (:559-934) / ForceBeamColumn3d::update
  int update(const double* ug) {
    const int nb = NB(), ns = NS(), nsec = fe.nsec;
    double v[6] = {0}, dv[6] = {0};
    fe.basic_disp(ug, v);
    double nrm = 0.0;
    for (int k = 0; k < nb; k++) {
      dv[k] = v[k] - vprev[k];
      nrm += dv[k] * dv[k];
    }
    for (int k = 0; k < nb; k++) vprev[k] = v[k];
    if (initialFlag != 0 && std::sqrt(nrm) <= DBL_EPSILON) return 0;
    double vin[6];
    for (int k = 0; k < nb; k++) vin[k] = v[k] - dv[k];
    const double L = fe.L;
    const int maxSubdivisions = fe.is3d ? 10 : 4;
    const double factor = 10.0;
    int numSubdivide = 1;
    bool converged = false;
    double dvToDo[6], dvTrial[6];
    for (int k = 0; k < nb; k++) dvToDo[k] = dvTrial[k] = dv[k];
    std::vector<std::array<double, 4>> vsSub(nsec), SsrSub(nsec);
    std::vector<std::array<double, 16>> fsSub(nsec);
    while (!converged && numSubdivide <= maxSubdivisions) {
      for (int l = 0; l < 3; l++) {
        double SeTrial[6], kvTrial[36];
        for (int k = 0; k < nb; k++) SeTrial[k] = Se[k];
        for (int k = 0; k < nb * nb; k++) kvTrial[k] = kv[k];
        vsSub = vs;
        fsSub = fs;
        SsrSub = Ssr;
        for (int r = 0; r < nb; r++) {
          double a = 0.0;
          for (int c = 0; c < nb; c++) a += kvTrial[r * nb + c] * dvTrial[c];
          SeTrial[r] += a;
        }
        const int numIters = l == 1 ? 10 * maxIters : maxIters;
        for (int j = 0; j < numIters; j++) {
          double f[36] = {0}, vr[6] = {0};
          for (int i = 0; i < nsec; i++) {
            const double xL = xi[i], xL1 = xL - 1.0, wtL = wt[i] * L;
            double Ss[4] = {0}, dSs[4], dvs[4];
            Ss[0] = SeTrial[0];
            Ss[1] = xL1 * SeTrial[1] + xL * SeTrial[2];
            if (fe.is3d) {
              Ss[2] = xL1 * SeTrial[3] + xL * SeTrial[4];
              Ss[3] = SeTrial[5];
            }
            for (int a = 0; a < ns; a++) dSs[a] = Ss[a] - SsrSub[i][a];
            double f0[16];
            const double* fsel = fsSub[i].data();
            if (l == 1 || (l == 2 && j == 0)) {
              initial_flexibility(i, f0);
              fsel = f0;
            }
            for (int a = 0; a < ns; a++) {
              double acc = 0.0;
              for (int b = 0; b < ns; b++) acc += fsel[a * ns + b] * dSs[b];
              dvs[a] = acc;
            }
            if (initialFlag != 0)
              for (int a = 0; a < ns; a++) vsSub[i][a] += dvs[a];
            fe.secs[i].set_trial(vsSub[i].data());
            for (int a = 0; a < ns; a++) SsrSub[i][a] = fe.secs[i].s[a];
            double ks[16];
            section_ks(i, ks);
            small_inverse(ns, ks, fsSub[i].data());
            for (int a = 0; a < ns; a++) dSs[a] = Ss[a] - SsrSub[i][a];
            for (int a = 0; a < ns; a++) {
              double acc = 0.0;
              for (int b = 0; b < ns; b++) acc += fsSub[i][a * ns + b] * dSs[b];
              dvs[a] = acc;
            }
            // f += b^T fs b wtL through fb (ns x nb), section codes P, MZ, MY, T
            const double* fSec = fsSub[i].data();
            double fb[4][6] = {{0}};
            for (int ii = 0; ii < ns; ii++) {
              for (int jj = 0; jj < ns; jj++) {
                const double tmp = fSec[jj * ns + ii] * wtL;
                if (ii == 0) fb[jj][0] += tmp;
                else if (ii == 1) {
                  fb[jj][1] += xL1 * tmp;
                  fb[jj][2] += xL * tmp;
                } else if (ii == 2) {
                  fb[jj][3] += xL1 * tmp;
                  fb[jj][4] += xL * tmp;
                } else
                  fb[jj][5] += tmp;
              }
            }
            for (int ii = 0; ii < ns; ii++)
              for (int jj = 0; jj < nb; jj++) {
                const double tmp = fb[ii][jj];
                if (ii == 0) f[0 * nb + jj] += tmp;
                else if (ii == 1) {
                  f[1 * nb + jj] += xL1 * tmp;
                  f[2 * nb + jj] += xL * tmp;
                } else if (ii == 2) {
                  f[3 * nb + jj] += xL1 * tmp;
                  f[4 * nb + jj] += xL * tmp;
                } else
                  f[5 * nb + jj] += tmp;
              }
            // vr += b^T (vs + dvs) wtL
            for (int ii = 0; ii < ns; ii++) {
              const double dei = (dvs[ii] + vsSub[i][ii]) * wtL;
              if (ii == 0) vr[0] += dei;
              else if (ii == 1) {
                vr[1] += xL1 * dei;
                vr[2] += xL * dei;
              } else if (ii == 2) {
                vr[3] += xL1 * dei;
                vr[4] += xL * dei;
              } else
                vr[5] += dei;
            }
          }
          small_inverse(nb, f, kvTrial);
          double dvr[6], dSe[6], dW = 0.0;
          for (int k = 0; k < nb; k++) dvr[k] = vin[k] + dvTrial[k] - vr[k];
          for (int r = 0; r < nb; r++) {
            double a = 0.0;
            for (int c = 0; c < nb; c++) a += kvTrial[r * nb + c] * dvr[c];
            dSe[r] = a;
          }
          for (int k = 0; k < nb; k++) dW += dvr[k] * dSe[k];
          for (int k = 0; k < nb; k++) SeTrial[k] += dSe[k];
          if (std::fabs(dW) < tol) {
            double n2 = 0.0;
            for (int k = 0; k < nb; k++) {
              dvToDo[k] -= dvTrial[k];
              vin[k] += dvTrial[k];
              n2 += dvToDo[k] * dvToDo[k];
            }
            if (std::sqrt(n2) <= DBL_EPSILON) converged = true;
            else {
              for (int k = 0; k < nb; k++) dvTrial[k] = dvToDo[k];
              numSubdivide = 1;
            }
            for (int k = 0; k < nb * nb; k++) kv[k] = kvTrial[k];
            for (int k = 0; k < nb; k++) Se[k] = SeTrial[k];
            vs = vsSub;
            fs = fsSub;
            Ssr = SsrSub;
            j = numIters + 1;
            l = 3;
          } else if (j == numIters - 1 && l == 2) {
            for (int k = 0; k < nb; k++) dvTrial[k] /= factor;
            numSubdivide++;
          }
        }
      }
    }
    if (!converged) return -1;
    initialFlag = 1;
    return 0;
  }

  void tangent(double* K) const {
    double kb[6][6] = {{0}};
    const int nb = NB();
    for (int r = 0; r < nb; r++)
      for (int c = 0; c < nb; c++) kb[r][c] = kv[r * nb + c];
    fe.global_stiff(kb, K);
  }
  void resisting_force(double* P) const {
    double qb[6] = {0};
    for (int k = 0; k < NB(); k++) qb[k] = Se[k];
    fe.global_force(qb, P);
  }
  // ForceBeamColumn2d::commitState (:276-308)
  void commit() {
    for (int i = 0; i < fe.nsec; i++) vscommit[i] = vs[i];
    fe.commit();
    for (int k = 0; k < 36; k++) kvcommit[k] = kv[k];
    for (int k = 0; k < 6; k++) {
      Secommit[k] = Se[k];
      vprev_c[k] = vprev[k];
    }
  }
  // ForceBeamColumn2d::revertToLastCommit (:310-342)
  void revert() {
    const int ns = NS();
    for (int i = 0; i < fe.nsec; i++) {
      vs[i] = vscommit[i];
      fe.secs[i].revert();
      fe.secs[i].set_trial(vs[i].data());
      for (int a = 0; a < ns; a++) Ssr[i][a] = fe.secs[i].s[a];
      double ks[16];
      section_ks(i, ks);
      small_inverse(ns, ks, fs[i].data());
    }
    for (int k = 0; k < 36; k++) kv[k] = kvcommit[k];
    for (int k = 0; k < 6; k++) {
      Se[k] = Secommit[k];
      vprev[k] = vprev_c[k];
    }
    initialFlag = 0;
  }
  // ForceBeamColumn2d::revertToStart (:344-376)
  void revert_to_start() {
    for (int i = 0; i < fe.nsec; i++) {
      fs[i].fill(0.0);
      vs[i].fill(0.0);
      Ssr[i].fill(0.0);
      vscommit[i].fill(0.0);
    }
    fe.revert_to_start();
    for (int k = 0; k < 36; k++) kv[k] = kvcommit[k] = 0.0;
    for (int k = 0; k < 6; k++) Se[k] = Secommit[k] = vprev[k] = vprev_c[k] = 0.0;
    initialFlag = 0;
  }
};

}  // namespace orc
