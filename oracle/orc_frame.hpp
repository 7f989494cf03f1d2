// Oracle: CPU restatement of the frame path -- FiberSection2d/3d over
// Steel02 / Concrete02 / Elastic fibers, LinearCrdTransf2d/3d and
// DispBeamColumn2d/3d with FrameQuadrature<GaussLegendre<1,5>>.
// TEST INFRASTRUCTURE ONLY (the checker for the GPU library).
#pragma once
#include <vector>

#include "../include/xara_gpu.h"
#include "orc_materials.hpp"

namespace orc {

// GaussLegendre<1,5>, SRC/quadrature/GaussLegendre1D.hpp:74-92 (the middle
// weight is the truncated literal 0.568888888888889)
static const double kGL5pts[5] = {-0.90617984593866396, -0.53846931010568311, 0.0, 0.53846931010568311,
                                  0.90617984593866396};
static const double kGL5wts[5] = {0.23692688505618942, 0.47862867049936619, 0.568888888888889,
                                  0.47862867049936619, 0.23692688505618942};

struct FiberDef {
  int kind = 0;  // 1 steel02, 2 concrete02, 3 elastic
  int mat_tag = 0;
  double y = 0, z = 0, A = 0;
  Steel02Params steel{};
  Concrete02Params conc{};
  double E = 0;
};
struct FiberSectionDef {
  bool is3d = false;
  double GJ = -1.0;
  bool compute_centroid = true;
  std::vector<FiberDef> fibers;
};

struct FiberState {
  Steel02State st;
  Concrete02State cs;
  double eps = 0, epsP = 0;  // elastic fiber
};

// UniaxialMaterial::setTrial (UniaxialMaterial.cpp:70-81): setTrialStrain,
// then getStress / getTangent.
inline int fiber_set_trial(const FiberDef& f, FiberState& s, double strain, double& stress,
                           double& tangent) {
  if (f.kind == 1) {
    int r = steel02_set_trial_strain(f.steel, s.st, strain);
    stress = s.st.sig;
    tangent = s.st.e;
    return r;
  } else if (f.kind == 2) {
    int r = concrete02_set_trial_strain(f.conc, s.cs, strain);
    stress = s.cs.sig;
    tangent = s.cs.e;
    return r;
  }
  s.eps = strain;
  stress = f.E * strain;
  tangent = f.E;
  return 0;
}

// A fiber section instance (one per integration point).
struct FiberSection {
  const FiberSectionDef* def = nullptr;
  double yBar = 0, zBar = 0;
  std::vector<FiberState> fs;
  double e[4] = {0, 0, 0, 0};
  double s[4] = {0, 0, 0, 0};
  double k[16] = {0};  // column-major order x order (kData)
  int order() const { return def->is3d ? (def->GJ >= 0 ? 4 : 3) : 2; }

  void init(const FiberSectionDef* d) {
    def = d;
    fs.assign(d->fibers.size(), FiberState{});
    // centroid recomputed at every addFiber (FiberSection2d.cpp:151-156,
    // FiberSection3d.cpp:338-345)
    double ABar = 0, QzBar = 0, QyBar = 0;
    yBar = zBar = 0;
    if (d->compute_centroid) {
      for (auto& f : d->fibers) {
        ABar += f.A;
        QzBar += f.y * f.A;
        QyBar += f.z * f.A;
        yBar = QzBar / ABar;
        zBar = QyBar / ABar;
      }
    }
    revert_to_start();
  }
  void revert_to_start() {
    for (size_t i = 0; i < fs.size(); i++) {
      const FiberDef& f = def->fibers[i];
      if (f.kind == 1) steel02_revert_to_start(f.steel, fs[i].st);
      else if (f.kind == 2) concrete02_revert_to_start(f.conc, fs[i].cs);
      fs[i].eps = fs[i].epsP = 0;
    }
  }

  // FiberSection2d::setTrialSectionDeformation (FiberSection2d.cpp:225-262)
  // FiberSection3d::setTrialSectionDeformation (FiberSection3d.cpp:423-476)
  int set_trial(const double* dfm) {
    int ord = order();
    for (int i = 0; i < ord; i++) e[i] = dfm[i];
    for (int i = 0; i < 16; i++) k[i] = 0.0;
    for (int i = 0; i < 4; i++) s[i] = 0.0;
    int res = 0;
    if (!def->is3d) {
      const double d0 = dfm[0], d1 = dfm[1];
      for (size_t i = 0; i < fs.size(); i++) {
        const FiberDef& f = def->fibers[i];
        const double y = f.y - yBar;
        const double A = f.A;
        double strain = d0 - y * d1;
        double tangent, stress;
        res += fiber_set_trial(f, fs[i], strain, stress, tangent);
        double ks0 = tangent * A;
        double ks1 = ks0 * -y;
        k[0] += ks0;
        k[1] += ks1;
        k[3] += ks1 * -y;
        double fs0 = stress * A;
        s[0] += fs0;
        s[1] += fs0 * -y;
      }
      k[2] = k[1];
    } else {
      // 3D kData is 4x4 column-major regardless of torsion
      const double e0 = dfm[0], e1 = dfm[1], e2 = dfm[2];
      for (size_t i = 0; i < fs.size(); i++) {
        const FiberDef& f = def->fibers[i];
        const double y = f.y - yBar;
        const double z = f.z - zBar;
        const double A = f.A;
        double strain = e0 - y * e1 + z * e2;
        double tangent, stress;
        res += fiber_set_trial(f, fs[i], strain, stress, tangent);
        double EA = tangent * A;
        k[0] += EA;
        k[1] += -y * EA;
        k[2] += z * EA;
        k[5] += y * y * EA;
        k[10] += z * z * EA;
        k[6] += -y * z * EA;
        double fs0 = stress * A;
        s[0] += fs0;
        s[1] += -y * fs0;
        s[2] += z * fs0;
      }
      k[4] = k[1];
      k[8] = k[2];
      k[9] = k[6];
      if (def->GJ >= 0) {
        // torsion ElasticMaterial(GJ): setTrial(e3)
        s[3] = def->GJ * dfm[3];
        k[15] = def->GJ;
      }
    }
    return res;
  }
  // ks(i,j) for the section's order (column-major storage; 3D uses ld=4)
  double ks(int i, int j) const {
    int ld = def->is3d ? 4 : 2;
    return k[j * ld + i];
  }
  void commit() {
    for (size_t i = 0; i < fs.size(); i++) {
      const FiberDef& f = def->fibers[i];
      if (f.kind == 1) steel02_commit(fs[i].st);
      else if (f.kind == 2) concrete02_commit(fs[i].cs);
      else fs[i].epsP = fs[i].eps;
    }
  }
  // FiberSection2d::revertToLastCommit (:376-409) keeps the reference's
  // assignment `sData[0] = fs0` (not +=) at :401.
  // FiberSection3d::revertToLastCommit (:613-650) does not zero kData[10].
  void revert() {
    if (!def->is3d) {
      k[0] = k[1] = k[2] = k[3] = 0.0;
      s[0] = s[1] = 0.0;
      for (size_t i = 0; i < fs.size(); i++) {
        const FiberDef& f = def->fibers[i];
        const double y = f.y - yBar;
        const double A = f.A;
        double tangent, stress;
        revert_fiber(f, fs[i], stress, tangent);
        double ks0 = tangent * A;
        double ks1 = ks0 * -y;
        k[0] += ks0;
        k[1] += ks1;
        k[3] += ks1 * -y;
        double fs0 = stress * A;
        s[0] = fs0;
        s[1] = fs0 * -y;
      }
      k[2] = k[1];
    } else {
      for (int i = 0; i <= 8; i++) k[i] = 0.0;
      k[15] = 0.0;
      s[0] = s[1] = s[2] = s[3] = 0.0;
      for (size_t i = 0; i < fs.size(); i++) {
        const FiberDef& f = def->fibers[i];
        const double y = f.y - yBar;
        const double z = f.z - zBar;
        const double A = f.A;
        double tangent, stress;
        revert_fiber(f, fs[i], stress, tangent);
        double value = tangent * A;
        double vas1 = -y * value;
        double vas2 = z * value;
        double vas1as2 = vas1 * z;
        k[0] += value;
        k[1] += vas1;
        k[2] += vas2;
        k[5] += vas1 * -y;
        k[6] += vas1as2;
        k[10] += vas2 * z;
        double fs0 = stress * A;
        s[0] += fs0;
        s[1] += fs0 * -y;
        s[2] += fs0 * z;
      }
      k[4] = k[1];
      k[8] = k[2];
      k[9] = k[6];
      if (def->GJ >= 0) {
        k[15] = def->GJ;
        // torsion stress is not restored by the reference revert
      }
    }
  }
  static void revert_fiber(const FiberDef& f, FiberState& st, double& stress, double& tangent) {
    if (f.kind == 1) {
      steel02_revert(st.st);
      stress = st.st.sig;
      tangent = st.st.e;
    } else if (f.kind == 2) {
      concrete02_revert(st.cs);
      stress = st.cs.sig;
      tangent = st.cs.e;
    } else {
      st.eps = st.epsP;
      stress = f.E * st.eps;
      tangent = f.E;
    }
  }
};

// DispBeamColumn2d/3d with LinearCrdTransf2d/3d (no offsets, no initial
// displacements, no element loads: q0 = p0 = 0).
struct FrameElement {
  bool is3d = false;
  int nsec = 5;
  double L = 0, cosT = 0, sinT = 0;
  double R[3][3] = {{0}};
  double rho = 0;
  std::vector<FiberSection> secs;
  double q[6] = {0};

  int ndof() const { return is3d ? 12 : 6; }

  int init(bool three_d, const double ci[3], const double cj[3], const double* vecxz,
           const FiberSectionDef& sd, int nip, double rho_) {
    is3d = three_d;
    nsec = nip;
    rho = rho_;
    if (!is3d) {
      // LinearCrdTransf2d::computeElemtLengthAndOrient, :175-221
      double dx0 = cj[0] - ci[0], dx1 = cj[1] - ci[1];
      L = std::sqrt(dx0 * dx0 + dx1 * dx1);
      if (L == 0.0) return -2;
      cosT = dx0 / L;
      sinT = dx1 / L;
    } else {
      // LinearCrdTransf3d::computeElemtLengthAndOrient (:203) + getLocalAxes
      double dx[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
      L = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      if (L == 0.0) return -2;
      R[0][0] = dx[0] / L;
      R[0][1] = dx[1] / L;
      R[0][2] = dx[2] / L;
      double v[3] = {vecxz[0], vecxz[1], vecxz[2]};
      double x[3] = {R[0][0], R[0][1], R[0][2]};
      double y[3] = {v[1] * x[2] - v[2] * x[1], v[2] * x[0] - v[0] * x[2], v[0] * x[1] - v[1] * x[0]};
      double yn = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
      if (yn == 0) return -3;
      for (int i = 0; i < 3; i++) y[i] /= yn;
      double z[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
      for (int i = 0; i < 3; i++) {
        R[1][i] = y[i];
        R[2][i] = z[i];
      }
    }
    secs.assign(nsec, FiberSection{});
    for (auto& s : secs) s.init(&sd);
    return 0;
  }

  // LinearCrdTransf2d::getBasicTrialDisp (:248-298), 3d (:345-420)
  void basic_disp(const double* ug, double* ub) const {
    double oneOverL = 1.0 / L;
    if (!is3d) {
      double sl = sinT * oneOverL, cl = cosT * oneOverL;
      ub[0] = -cosT * ug[0] - sinT * ug[1] + cosT * ug[3] + sinT * ug[4];
      ub[1] = -sl * ug[0] + cl * ug[1] + ug[2] + sl * ug[3] - cl * ug[4];
      ub[2] = ub[1] + ug[5] - ug[2];
    } else {
      double ul[12];
      for (int b = 0; b < 4; b++)
        for (int r = 0; r < 3; r++)
          ul[3 * b + r] = R[r][0] * ug[3 * b] + R[r][1] * ug[3 * b + 1] + R[r][2] * ug[3 * b + 2];
      ub[0] = ul[6] - ul[0];
      double tmp = oneOverL * (ul[1] - ul[7]);
      ub[1] = ul[5] + tmp;
      ub[2] = ul[11] + tmp;
      tmp = oneOverL * (ul[8] - ul[2]);
      ub[3] = ul[4] + tmp;
      ub[4] = ul[10] + tmp;
      ub[5] = ul[9] - ul[3];
    }
  }

  // DispBeamColumn2d::update (:235-284) / DispBeamColumn3d::update (:235-292)
  int update(const double* ug) {
    double v[6];
    basic_disp(ug, v);
    double oneOverL = 1.0 / L;
    int err = 0;
    for (int i = 0; i < nsec; i++) {
      double xi = 0.5 * (1.0 + kGL5pts[i]);
      double xi6 = 6.0 * xi;
      double e[4];
      e[0] = oneOverL * v[0];
      e[1] = oneOverL * ((xi6 - 4.0) * v[1] + (xi6 - 2.0) * v[2]);
      if (is3d) {
        e[2] = oneOverL * ((xi6 - 4.0) * v[3] + (xi6 - 2.0) * v[4]);
        e[3] = oneOverL * v[5];
      }
      err += secs[i].set_trial(e);
    }
    return err;
  }

  // codes: 2D {P, MZ}; 3D {P, MZ, MY, T}
  static int code(bool three_d, int j) {
    (void)three_d;
    return j;  // 0 P, 1 MZ, 2 MY, 3 T
  }

  // getBasicStiff (2D :287-355) and the kb loop of 3D getTangentStiff (:407-537)
  void basic_stiff(double kb[6][6]) const {
    int nb = is3d ? 6 : 3;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) kb[i][j] = 0.0;
    double oneOverL = 1.0 / L;
    for (int i = 0; i < nsec; i++) {
      const FiberSection& s = secs[i];
      int order = s.order();
      double ka[4][6] = {{0}};
      double xi6 = 6.0 * (0.5 * (1.0 + kGL5pts[i]));
      double wti = (kGL5wts[i] * 0.5) * oneOverL;
      for (int j = 0; j < order; j++) {
        int c = code(is3d, j);
        if (c == 0) {
          for (int k = 0; k < order; k++) ka[k][0] += s.ks(k, j) * wti;
        } else if (c == 1) {
          for (int k = 0; k < order; k++) {
            double tmp = s.ks(k, j) * wti;
            ka[k][1] += (xi6 - 4.0) * tmp;
            ka[k][2] += (xi6 - 2.0) * tmp;
          }
        } else if (c == 2) {
          for (int k = 0; k < order; k++) {
            double tmp = s.ks(k, j) * wti;
            ka[k][3] += (xi6 - 4.0) * tmp;
            ka[k][4] += (xi6 - 2.0) * tmp;
          }
        } else {
          for (int k = 0; k < order; k++) ka[k][5] += s.ks(k, j) * wti;
        }
      }
      for (int j = 0; j < order; j++) {
        int c = code(is3d, j);
        if (c == 0) {
          for (int k = 0; k < nb; k++) kb[0][k] += ka[j][k];
        } else if (c == 1) {
          for (int k = 0; k < nb; k++) {
            double tmp = ka[j][k];
            kb[1][k] += (xi6 - 4.0) * tmp;
            kb[2][k] += (xi6 - 2.0) * tmp;
          }
        } else if (c == 2) {
          for (int k = 0; k < nb; k++) {
            double tmp = ka[j][k];
            kb[3][k] += (xi6 - 4.0) * tmp;
            kb[4][k] += (xi6 - 2.0) * tmp;
          }
        } else {
          for (int k = 0; k < nb; k++) kb[5][k] += ka[j][k];
        }
      }
    }
  }

  // basic forces q = sum B^T s wt (getResistingForce, 2D :665-726)
  void basic_force(double qb[6]) const {
    for (int i = 0; i < 6; i++) qb[i] = 0.0;
    for (int i = 0; i < nsec; i++) {
      const FiberSection& s = secs[i];
      int order = s.order();
      double xi6 = 6.0 * (0.5 * (1.0 + kGL5pts[i]));
      double wt = kGL5wts[i] * 0.5;
      for (int j = 0; j < order; j++) {
        double si = s.s[j] * wt;
        int c = code(is3d, j);
        if (c == 0) qb[0] += si;
        else if (c == 1) {
          qb[1] += (xi6 - 4.0) * si;
          qb[2] += (xi6 - 2.0) * si;
        } else if (c == 2) {
          qb[3] += (xi6 - 4.0) * si;
          qb[4] += (xi6 - 2.0) * si;
        } else
          qb[5] += si;
      }
    }
  }

  // LinearCrdTransf2d::getGlobalStiffMatrix (:610-733) /
  // LinearCrdTransf3d::getGlobalStiffMatrix (:768) -- row-major K
  void tangent(double* K) const {
    double kb[6][6];
    basic_stiff(kb);
    global_stiff(kb, K);
  }
  // the transformation part of tangent(): K = T^T kb T
  void global_stiff(const double kb[6][6], double* K) const {
    double oneOverL = 1.0 / L;
    if (!is3d) {
      double tmp[3][6];
      double sl = sinT * oneOverL, cl = cosT * oneOverL;
      for (int r = 0; r < 3; r++) {
        tmp[r][0] = -cosT * kb[r][0] - sl * (kb[r][1] + kb[r][2]);
        tmp[r][1] = -sinT * kb[r][0] + cl * (kb[r][1] + kb[r][2]);
        tmp[r][2] = kb[r][1];
        tmp[r][3] = -tmp[r][0];
        tmp[r][4] = -tmp[r][1];
        tmp[r][5] = kb[r][2];
      }
      for (int m = 0; m < 6; m++) {
        K[0 * 6 + m] = -cosT * tmp[0][m] - sl * (tmp[1][m] + tmp[2][m]);
        K[1 * 6 + m] = -sinT * tmp[0][m] + cl * (tmp[1][m] + tmp[2][m]);
        K[2 * 6 + m] = tmp[1][m];
        K[3 * 6 + m] = -K[0 * 6 + m];
        K[4 * 6 + m] = -K[1 * 6 + m];
        K[5 * 6 + m] = tmp[2][m];
      }
      return;
    }
    double tmp[12][12], kl[12][12];
    for (int i = 0; i < 6; i++) {
      tmp[i][0] = -kb[i][0];
      tmp[i][1] = oneOverL * (kb[i][1] + kb[i][2]);
      tmp[i][2] = -oneOverL * (kb[i][3] + kb[i][4]);
      tmp[i][3] = -kb[i][5];
      tmp[i][4] = kb[i][3];
      tmp[i][5] = kb[i][1];
      tmp[i][6] = kb[i][0];
      tmp[i][7] = -tmp[i][1];
      tmp[i][8] = -tmp[i][2];
      tmp[i][9] = kb[i][5];
      tmp[i][10] = kb[i][4];
      tmp[i][11] = kb[i][2];
    }
    for (int i = 0; i < 12; i++) {
      kl[0][i] = -tmp[0][i];
      kl[1][i] = oneOverL * (tmp[1][i] + tmp[2][i]);
      kl[2][i] = -oneOverL * (tmp[3][i] + tmp[4][i]);
      kl[3][i] = -tmp[5][i];
      kl[4][i] = tmp[3][i];
      kl[5][i] = tmp[1][i];
      kl[6][i] = tmp[0][i];
      kl[7][i] = -kl[1][i];
      kl[8][i] = -kl[2][i];
      kl[9][i] = tmp[5][i];
      kl[10][i] = tmp[4][i];
      kl[11][i] = tmp[2][i];
    }
    for (int m = 0; m < 12; m++)
      for (int b = 0; b < 4; b++)
        for (int c = 0; c < 3; c++)
          tmp[m][3 * b + c] =
              kl[m][3 * b] * R[0][c] + kl[m][3 * b + 1] * R[1][c] + kl[m][3 * b + 2] * R[2][c];
    for (int m = 0; m < 12; m++)
      for (int b = 0; b < 4; b++)
        for (int c = 0; c < 3; c++)
          K[(3 * b + c) * 12 + m] =
              R[0][c] * tmp[3 * b][m] + R[1][c] * tmp[3 * b + 1][m] + R[2][c] * tmp[3 * b + 2][m];
  }

  // getResistingForce + getGlobalResistingForce (2D :473-525; 3D :700)
  void resisting_force(double* P) const {
    double qb[6];
    basic_force(qb);
    global_force(qb, P);
  }
  // the transformation part of resisting_force(): P = T^T qb
  void global_force(const double qb[6], double* P) const {
    double oneOverL = 1.0 / L;
    if (!is3d) {
      double q0 = qb[0], q1 = qb[1], q2 = qb[2];
      double V = oneOverL * (q1 + q2);
      double pl[6] = {-q0, V, q1, q0, -V, q2};
      P[0] = cosT * pl[0] - sinT * pl[1];
      P[1] = sinT * pl[0] + cosT * pl[1];
      P[3] = cosT * pl[3] - sinT * pl[4];
      P[4] = sinT * pl[3] + cosT * pl[4];
      P[2] = pl[2];
      P[5] = pl[5];
      return;
    }
    double pl[12];
    pl[0] = -qb[0];
    pl[1] = oneOverL * (qb[1] + qb[2]);
    pl[2] = -oneOverL * (qb[3] + qb[4]);
    pl[3] = -qb[5];
    pl[4] = qb[3];
    pl[5] = qb[1];
    pl[6] = qb[0];
    pl[7] = -pl[1];
    pl[8] = -pl[2];
    pl[9] = qb[5];
    pl[10] = qb[4];
    pl[11] = qb[2];
    for (int b = 0; b < 4; b++)
      for (int c = 0; c < 3; c++)
        P[3 * b + c] = R[0][c] * pl[3 * b] + R[1][c] * pl[3 * b + 1] + R[2][c] * pl[3 * b + 2];
  }

  // lumped translational mass 0.5*rho*L (DispBeamColumn2d::getMass, cMass = 0)
  void lumped_mass(double* M) const {
    int nd = ndof();
    for (int i = 0; i < nd * nd; i++) M[i] = 0.0;
    if (rho == 0.0) return;
    double m = 0.5 * rho * L;
    int per = is3d ? 6 : 3, ntr = is3d ? 3 : 2;
    for (int nnode = 0; nnode < 2; nnode++)
      for (int j = 0; j < ntr; j++) {
        int i = nnode * per + j;
        M[i * nd + i] = m;
      }
  }

  void commit() {
    for (auto& s : secs) s.commit();
  }
  void revert() {
    for (auto& s : secs) s.revert();
  }
  void revert_to_start() {
    for (auto& s : secs) s.revert_to_start();
  }
  int point_data(int what, double* out) const {
    int ord = secs.empty() ? 0 : secs[0].order();
    for (int i = 0; i < nsec; i++) {
      const FiberSection& s = secs[i];
      if (what == XG_OUT_STRESS)
        for (int j = 0; j < ord; j++) out[i * ord + j] = s.s[j];
      else if (what == XG_OUT_STRAIN)
        for (int j = 0; j < ord; j++) out[i * ord + j] = s.e[j];
      else if (what == XG_OUT_TANGENT)
        for (int a = 0; a < ord; a++)
          for (int b = 0; b < ord; b++) out[i * ord * ord + a * ord + b] = s.ks(a, b);
      else
        return -1;
    }
    return what == XG_OUT_TANGENT ? ord * ord : ord;
  }
};

}  // namespace orc
