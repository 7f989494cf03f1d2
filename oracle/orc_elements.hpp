// Oracle: CPU restatement of the reference continuum elements on the
// state-determination path (stdBrick, FourNodeQuad).  TEST INFRASTRUCTURE
// ONLY -- the checker for the GPU library, never shipped in its path.
#pragma once
#include "orc_materials.hpp"

namespace orc {

// shp3d, SRC/interpolate/shp3d.cpp:32-172
inline void shp3d(const double xn[3], double& xsj, double shp[4][8], const double xl[3][8]) {
  double ap1 = 1.0 + xn[0];
  double am1 = 1.0 - xn[0];
  double ap2 = 1.0 + xn[1];
  double am2 = 1.0 - xn[1];
  double ap3 = 1.0 + xn[2];
  double am3 = 1.0 - xn[2];
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
  xsj = xs[0][0] * ad[0][0] + xs[0][1] * ad[1][0] + xs[0][2] * ad[2][0];
  double rxsj = 1.0 / xsj;
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) xs[i][j] = ad[i][j] * rxsj;
  for (int k = 0; k < 8; k++) {
    double c1 = shp[0][k] * xs[0][0] + shp[1][k] * xs[1][0] + shp[2][k] * xs[2][0];
    double c2 = shp[0][k] * xs[0][1] + shp[1][k] * xs[1][1] + shp[2][k] * xs[2][1];
    double c3 = shp[0][k] * xs[0][2] + shp[1][k] * xs[1][2] + shp[2][k] * xs[2][2];
    shp[0][k] = c1;
    shp[1][k] = c2;
    shp[2][k] = c3;
  }
}

// Brick quadrature, Brick.cpp:55-62: sg = -+1/sqrt(3), wg = 1; point index
// count = 4i+2j+k with xi=sg[i], eta=sg[j], zeta=sg[k] (Brick.cpp:754-775)
inline void brick_gauss_point(int g, double xn[3]) {
  const double root3 = std::sqrt(3.0);
  const double one_over_root3 = 1.0 / root3;
  const double sg[2] = {-one_over_root3, one_over_root3};
  xn[0] = sg[(g >> 2) & 1];
  xn[1] = sg[(g >> 1) & 1];
  xn[2] = sg[g & 1];
}

// Brick::update, Brick.cpp:719-835: strain at every Gauss point from the
// total trial displacement; the material return code is ignored (:839).
// xl[3][8] nodal coordinates, ul[8][3] nodal trial displacements.
inline void brick_update(const J2Params& mp, double dt, J2Point pts[8], const double xl[3][8],
                         const double ul[8][3]) {
  for (int g = 0; g < 8; g++) {
    double xn[3], xsj, shp[4][8];
    brick_gauss_point(g, xn);
    shp3d(xn, xsj, shp, xl);
    double strain[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 8; j++) {
      double b00 = shp[0][j], b11 = shp[1][j], b22 = shp[2][j];
      double b30 = shp[1][j], b31 = shp[0][j];
      double b41 = shp[2][j], b42 = shp[1][j];
      double b50 = shp[2][j], b52 = shp[0][j];
      double ul0 = ul[j][0], ul1 = ul[j][1], ul2 = ul[j][2];
      strain[0] += b00 * ul0;
      strain[1] += b11 * ul1;
      strain[2] += b22 * ul2;
      strain[3] += b30 * ul0 + b31 * ul1;
      strain[4] += b41 * ul1 + b42 * ul2;
      strain[5] += b50 * ul0 + b52 * ul2;
    }
    j2_3d_set_trial_strain(mp, dt, pts[g], strain);
  }
}

// Brick::formResidAndTangent, Brick.cpp:844-1024 (row-major stiff[24][24]).
// tang_flag = 1 forms the tangent as well; body force b[3].
inline void brick_form_resid_and_tangent(const J2Point pts[8], const double xl[3][8], const double b[3],
                                         int tang_flag, double stiff[576], double resid[24]) {
  for (int i = 0; i < 576; i++) stiff[i] = 0.0;
  for (int i = 0; i < 24; i++) resid[i] = 0.0;
  double Shape[4][8][8], dvol[8];
  for (int g = 0; g < 8; g++) {
    double xn[3], xsj, shp[4][8];
    brick_gauss_point(g, xn);
    shp3d(xn, xsj, shp, xl);
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 8; q++) Shape[p][q][g] = shp[p][q];
    dvol[g] = 1.0 * xsj;
  }
  for (int g = 0; g < 8; g++) {
    double shp[4][8];
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 8; q++) shp[p][q] = Shape[p][q][g];
    double stress[6], dd[36];
    j2_3d_get_stress(pts[g], stress);
    for (int k = 0; k < 6; k++) stress[k] *= dvol[g];
    if (tang_flag == 1) {
      j2_3d_get_tangent(pts[g], dd);
      for (int k = 0; k < 36; k++) dd[k] *= dvol[g];
    }
    for (int j = 0; j < 8; j++) {
      double b00 = shp[0][j], b11 = shp[1][j], b22 = shp[2][j];
      double b30 = shp[1][j], b31 = shp[0][j];
      double b41 = shp[2][j], b42 = shp[1][j];
      double b50 = shp[2][j], b52 = shp[0][j];
      double r0 = b00 * stress[0] + b30 * stress[3] + b50 * stress[5];
      double r1 = b11 * stress[1] + b31 * stress[3] + b41 * stress[4];
      double r2 = b22 * stress[2] + b42 * stress[4] + b52 * stress[5];
      double rj[3] = {r0, r1, r2};
      for (int p = 0; p < 3; p++) {
        resid[3 * j + p] += rj[p];
        resid[3 * j + p] -= dvol[g] * b[p] * shp[3][j];
      }
      if (tang_flag == 1) {
        // B matrix of node J (computeB, Brick.cpp:1047-1076), 6x3
        double BJ[6][3] = {{0}};
        BJ[0][0] = shp[0][j]; BJ[1][1] = shp[1][j]; BJ[2][2] = shp[2][j];
        BJ[3][0] = shp[1][j]; BJ[3][1] = shp[0][j];
        BJ[4][1] = shp[2][j]; BJ[4][2] = shp[1][j];
        BJ[5][0] = shp[2][j]; BJ[5][2] = shp[0][j];
        // BJtranD = BJ^T dd   (3x6), Matrix::addMatrixProduct order
        double BJtranD[3][6];
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 6; q++) {
            double acc = 0.0;
            for (int k = 0; k < 6; k++) acc += BJ[k][p] * dd[k * 6 + q];
            BJtranD[p][q] = acc;
          }
        for (int kn = 0; kn < 8; kn++) {
          double BK[6][3] = {{0}};
          BK[0][0] = shp[0][kn]; BK[1][1] = shp[1][kn]; BK[2][2] = shp[2][kn];
          BK[3][0] = shp[1][kn]; BK[3][1] = shp[0][kn];
          BK[4][1] = shp[2][kn]; BK[4][2] = shp[1][kn];
          BK[5][0] = shp[2][kn]; BK[5][2] = shp[0][kn];
          for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
              double acc = 0.0;
              for (int k = 0; k < 6; k++) acc += BJtranD[p][k] * BK[k][q];
              stiff[(3 * j + p) * 24 + 3 * kn + q] += acc;
            }
        }
      }
    }
  }
}

// FourNodeQuad quadrature GaussLegendre<2,4>, GaussLegendre2D.hpp:23-38
static const double kQuadPts[4][2] = {{-0.57735026918962573, -0.57735026918962573},
                                      {-0.57735026918962573, 0.57735026918962573},
                                      {0.57735026918962573, -0.57735026918962573},
                                      {0.57735026918962573, 0.57735026918962573}};
static const double kQuadWts[4] = {1.0, 1.0, 1.0, 1.0};

// FourNodeQuad::shapeFunction, FourNodeQuad.cpp:1128-1195.  crd[4][2].
inline double quad_shape(double xi, double eta, const double crd[4][2], double shp[3][4]) {
  double oneMinuseta = 1.0 - eta, onePluseta = 1.0 + eta;
  double oneMinusxi = 1.0 - xi, onePlusxi = 1.0 + xi;
  shp[2][0] = 0.25 * oneMinusxi * oneMinuseta;
  shp[2][1] = 0.25 * onePlusxi * oneMinuseta;
  shp[2][2] = 0.25 * onePlusxi * onePluseta;
  shp[2][3] = 0.25 * oneMinusxi * onePluseta;
  double J[2][2];
  J[0][0] = 0.25 * (-crd[0][0] * oneMinuseta + crd[1][0] * oneMinuseta + crd[2][0] * (onePluseta) -
                    crd[3][0] * (onePluseta));
  J[0][1] = 0.25 * (-crd[0][0] * oneMinusxi - crd[1][0] * onePlusxi + crd[2][0] * onePlusxi +
                    crd[3][0] * oneMinusxi);
  J[1][0] = 0.25 * (-crd[0][1] * oneMinuseta + crd[1][1] * oneMinuseta + crd[2][1] * onePluseta -
                    crd[3][1] * onePluseta);
  J[1][1] = 0.25 * (-crd[0][1] * oneMinusxi - crd[1][1] * onePlusxi + crd[2][1] * onePlusxi +
                    crd[3][1] * oneMinusxi);
  double detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  double oneOverdetJ = 1.0 / detJ;
  double L[2][2];
  L[0][0] = J[1][1] * oneOverdetJ;
  L[1][0] = -J[0][1] * oneOverdetJ;
  L[0][1] = -J[1][0] * oneOverdetJ;
  L[1][1] = J[0][0] * oneOverdetJ;
  double L00 = 0.25 * L[0][0], L10 = 0.25 * L[1][0], L01 = 0.25 * L[0][1], L11 = 0.25 * L[1][1];
  double L00oneMinuseta = L00 * oneMinuseta, L00onePluseta = L00 * onePluseta;
  double L01oneMinusxi = L01 * oneMinusxi, L01onePlusxi = L01 * onePlusxi;
  double L10oneMinuseta = L10 * oneMinuseta, L10onePluseta = L10 * onePluseta;
  double L11oneMinusxi = L11 * oneMinusxi, L11onePlusxi = L11 * onePlusxi;
  shp[0][0] = -L00oneMinuseta - L01oneMinusxi;
  shp[0][1] = L00oneMinuseta - L01onePlusxi;
  shp[0][2] = L00onePluseta + L01onePlusxi;
  shp[0][3] = -L00onePluseta + L01oneMinusxi;
  shp[1][0] = -L10oneMinuseta - L11oneMinusxi;
  shp[1][1] = L10oneMinuseta - L11onePlusxi;
  shp[1][2] = L10onePluseta + L11onePlusxi;
  shp[1][3] = -L10onePluseta + L11oneMinusxi;
  return detJ;
}

// FourNodeQuad::update, FourNodeQuad.cpp:190-223 (u[4][2])
inline void quad_update(const J2Params& mp, double dt, J2Point pts[4], const double crd[4][2],
                        const double u[4][2]) {
  for (int i = 0; i < 4; i++) {
    double shp[3][4];
    quad_shape(kQuadPts[i][0], kQuadPts[i][1], crd, shp);
    double eps[3] = {0, 0, 0};
    for (int beta = 0; beta < 4; beta++) {
      eps[0] += shp[0][beta] * u[beta][0];
      eps[1] += shp[1][beta] * u[beta][1];
      eps[2] += shp[0][beta] * u[beta][1] + shp[1][beta] * u[beta][0];
    }
    j2_ps_set_trial_strain(mp, dt, pts[i], eps);
  }
}

// FourNodeQuad::getTangentStiff, FourNodeQuad.cpp:226-281 (row-major 8x8)
inline void quad_tangent(const J2Point pts[4], const double crd[4][2], double thk, double K[64]) {
  for (int i = 0; i < 64; i++) K[i] = 0.0;
  for (int i = 0; i < 4; i++) {
    double shp[3][4];
    double dvol = quad_shape(kQuadPts[i][0], kQuadPts[i][1], crd, shp);
    dvol *= (thk * kQuadWts[i]);
    double D[9];
    j2_ps_get_tangent(pts[i], D);
    const double D00 = D[0], D01 = D[1], D02 = D[2], D10 = D[3], D11 = D[4], D12 = D[5], D20 = D[6],
                 D21 = D[7], D22 = D[8];
    for (int alpha = 0, ia = 0; alpha < 4; alpha++, ia += 2) {
      for (int beta = 0, ib = 0; beta < 4; beta++, ib += 2) {
        double DB[3][2];
        DB[0][0] = dvol * (D00 * shp[0][beta] + D02 * shp[1][beta]);
        DB[1][0] = dvol * (D10 * shp[0][beta] + D12 * shp[1][beta]);
        DB[2][0] = dvol * (D20 * shp[0][beta] + D22 * shp[1][beta]);
        DB[0][1] = dvol * (D01 * shp[1][beta] + D02 * shp[0][beta]);
        DB[1][1] = dvol * (D11 * shp[1][beta] + D12 * shp[0][beta]);
        DB[2][1] = dvol * (D21 * shp[1][beta] + D22 * shp[0][beta]);
        K[ia * 8 + ib] += shp[0][alpha] * DB[0][0] + shp[1][alpha] * DB[2][0];
        K[ia * 8 + ib + 1] += shp[0][alpha] * DB[0][1] + shp[1][alpha] * DB[2][1];
        K[(ia + 1) * 8 + ib] += shp[1][alpha] * DB[1][0] + shp[0][alpha] * DB[2][0];
        K[(ia + 1) * 8 + ib + 1] += shp[1][alpha] * DB[1][1] + shp[0][alpha] * DB[2][1];
      }
    }
  }
}

// FourNodeQuad::getResistingForce, FourNodeQuad.cpp:507-553.  The consistent
// pressure load (setPressureLoadAtNodes) is not used on the path (pressure 0).
inline void quad_resisting_force(const J2Point pts[4], const double crd[4][2], double thk,
                                 const double b[2], double P[8]) {
  for (int i = 0; i < 8; i++) P[i] = 0.0;
  for (int i = 0; i < 4; i++) {
    double shp[3][4];
    double dvol = quad_shape(kQuadPts[i][0], kQuadPts[i][1], crd, shp);
    dvol *= (thk * kQuadWts[i]);
    double sigma[3];
    j2_ps_get_stress(pts[i], sigma);
    for (int alpha = 0, ia = 0; alpha < 4; alpha++, ia += 2) {
      P[ia] += dvol * (shp[0][alpha] * sigma[0] + shp[1][alpha] * sigma[2]);
      P[ia + 1] += dvol * (shp[1][alpha] * sigma[1] + shp[0][alpha] * sigma[2]);
      P[ia] -= dvol * (shp[2][alpha] * b[0]);
      P[ia + 1] -= dvol * (shp[2][alpha] * b[1]);
    }
  }
}

}  // namespace orc
