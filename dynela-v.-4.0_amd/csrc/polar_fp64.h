// polar_fp64.h -- FP64 device restatement of the incremental logarithmic
// strain and rotation of one integration point.
//
// Reference: Tensor2::polarCuppenLnU (Tensor2.C:1442-1457) = buildFTF
// (Tensor2.C:1315-1327) + dsyevd3 (Eigen3x3.C:454-634, Kopp's hybrid
// Householder / Cuppen divide-and-conquer method, arXiv:physics/0610206) +
// polarExtractLnU (Tensor2.C:1237-1312).  The arithmetic follows the same
// steps so the FP64 engine tracks the FP64 oracle to rounding; the secular
// solver is a template on the index permutation so every array index is a
// compile-time constant (no scratch memory on the GPU).
#pragma once

#include "engine.h"

namespace dnlg {

__device__ __forceinline__ void eig_tridiag(const double A[3][3], double Q[3][3], double d[3], double e[2]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Q[i][j] = (i == j) ? 1.0 : 0.0;
  const double h = sqd(A[0][1]) + sqd(A[0][2]);
  const double g = (A[0][1] > 0) ? -sqrt(h) : sqrt(h);
  e[0] = g;
  double f = g * A[0][1];
  double u1 = A[0][1] - g, u2 = A[0][2];
  double omega = h - f;
  if (omega > 0.0) {
    omega = 1.0 / omega;
    const double f1 = A[1][1] * u1 + A[1][2] * u2;
    const double f2 = A[1][2] * u1 + A[2][2] * u2;
    double q1 = omega * f1, q2 = omega * f2;
    double K = 0.0;
    K += u1 * f1;
    K += u2 * f2;
    K *= 0.5 * sqd(omega);
    q1 = q1 - K * u1;
    q2 = q2 - K * u2;
    d[0] = A[0][0];
    d[1] = A[1][1] - 2.0 * q1 * u1;
    d[2] = A[2][2] - 2.0 * q2 * u2;
    f = omega * u1;
    Q[1][1] = Q[1][1] - f * u1;
    Q[2][1] = Q[2][1] - f * u2;
    f = omega * u2;
    Q[1][2] = Q[1][2] - f * u1;
    Q[2][2] = Q[2][2] - f * u2;
    e[1] = A[1][2] - q1 * u2 - u1 * q2;
  } else {
    d[0] = A[0][0];
    d[1] = A[1][1];
    d[2] = A[2][2];
    e[1] = A[1][2];
  }
}

__device__ __forceinline__ void eig_2x2(double a, double b, double c, double &rt1, double &rt2, double &cs,
                                        double &sn) {
  const double sm = a + c, df = a - c;
  const double rt = sqrt(sqd(df) + 4.0 * b * b);
  double t;
  if (sm > 0.0) {
    rt1 = 0.5 * (sm + rt);
    t = 1.0 / rt1;
    rt2 = (a * t) * c - (b * t) * b;
  } else if (sm < 0.0) {
    rt2 = 0.5 * (sm - rt);
    t = 1.0 / rt2;
    rt1 = (a * t) * c - (b * t) * b;
  } else {
    rt1 = 0.5 * rt;
    rt2 = -0.5 * rt;
  }
  cs = (df > 0.0) ? df + rt : df - rt;
  if (fabs(cs) > 2.0 * fabs(b)) {
    t = -2.0 * b / cs;
    sn = 1.0 / sqrt(1.0 + sqd(t));
    cs = t * sn;
  } else if (fabs(b) == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else {
    t = -0.5 * cs / b;
    cs = 1.0 / sqrt(1.0 + sqd(t));
    sn = t * cs;
  }
  if (df > 0.0) {
    t = cs;
    cs = -sn;
    sn = t;
  }
}

// Roots w of 1 + sum z_j / (d_j - w) = 0 (Eigen3x3.C, the secular step of
// dsyevd3) and P[j][i] = d_j - w_i, each root refined by the reference's
// safeguarded Newton / bisection iteration and stopping test from a closed-
// form starting value.  One deliberate difference: the starting values are
// the roots of the characteristic cubic written in d - m, m the mean of d,
// where the reference writes it in d.  For F^T F near the identity (every
// strain increment) the reference's coefficients cancel to a few digits,
// its starting values miss by about 1 % of the eigenvalue gaps, and the
// refinement takes about 9 iterations per root, many of them bisections;
// from the shifted cubic's values it takes 2, and lands inside the same
// acceptance band of the stopping test (same eigen pairs to rounding; the
// CPU restatement of the reference keeps the reference's start).
template <int I0, int I1, int I2>
__device__ __forceinline__ void eig_secular(const double d[3], const double z[3], double w[3], double P[3][3]) {
  double a[4];
  if (z[0] > 0) {
    a[0] = d[I0];
    a[1] = d[I1];
    a[2] = d[I2];
    a[3] = fabs(d[0] + 3.0 * z[0]) + fabs(d[1] + 3.0 * z[1]) + fabs(d[2] + 3.0 * z[2]);
  } else {
    a[0] = -fabs(d[0] + 3.0 * z[0]) - fabs(d[1] + 3.0 * z[1]) - fabs(d[2] + 3.0 * z[2]);
    a[1] = d[I0];
    a[2] = d[I1];
    a[3] = d[I2];
  }
  const double msh = (d[0] + d[1] + d[2]) * (1.0 / 3.0);
  const double e0 = d[0] - msh, e1 = d[1] - msh, e2 = d[2] - msh;
  const double t0 = e1 * e2, t1 = e0 * e2, t2 = e0 * e1;
  const double cg = t0 * e0 + (z[0] * t0 + z[1] * t1 + z[2] * t2);
  const double cb = (z[0] * (e1 + e2) + z[1] * (e0 + e2) + z[2] * (e0 + e1)) + (t0 + t1 + t2);
  const double ca = (z[0] + z[1] + z[2]) + (e0 + e1 + e2);
  const double p = sqd(ca) - 3.0 * cb;
  const double q = ca * (p - (3.0 / 2.0) * cb) + (27.0 / 2.0) * cg;
  const double sqrt_p = sqrt(fabs(p));
  double phi = 27.0 * (0.25 * sqd(cb) * (p - cb) - cg * (q - 27.0 / 4.0 * cg));
  phi = (1.0 / 3.0) * atan2(sqrt(fabs(phi)), q);
  double sphi, cphi;
  sincos(phi, &sphi, &cphi);
  const double c = sqrt_p * cphi;
  const double s = (1.0 / kSqrt3) * sqrt_p * fabs(sphi);
  double x0[3];
  x0[0] = x0[1] = x0[2] = (1.0 / 3.0) * (ca - c);
  if (c > s) {
    x0[0] -= s;
    x0[1] += s;
    x0[2] += c;
  } else if (c < -s) {
    x0[0] += c;
    x0[1] -= s;
    x0[2] += s;
  } else {
    x0[0] -= s;
    x0[1] += c;
    x0[2] += s;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) x0[i] += msh;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double xl = a[i], xh = a[i + 1];
    double dx = 0.5 * (xh - xl), dxold = dx;
    double dd0, dd1, dd2, x, delta;
    if (dx == 0.0) {
      w[i] = xl;
      P[0][i] = d[0] - xl;
      P[1][i] = d[1] - xl;
      P[2][i] = d[2] - xl;
      continue;
    }
    if (x0[i] >= xh) {
      delta = xh;
      x = -dx;
    } else if (x0[i] <= xl) {
      delta = xl;
      x = dx;
    } else {
      delta = x0[i];
      x = 0.0;
    }
    dd0 = d[0] - delta;
    dd1 = d[1] - delta;
    dd2 = d[2] - delta;
    double r0 = dd0 - x, r1 = dd1 - x, r2 = dd2 - x;
    xl -= delta;
    xh -= delta;
    if (z[0] < 0.0) {
      const double tmp = xh;
      xh = xl;
      xl = tmp;
    }
    for (int it = 0; it < 500; it++) {
      double F = 1.0, dF = 0.0, err = 1.0;
      // the three reciprocals from one division of their product (a few
      // ulp each; the stopping test allows 8 eps of slack), or three
      // divisions when the product leaves the normal range (a pole hit
      // exactly, or one close enough to underflow the product)
      double q0, q1, q2;
      const double r01 = r0 * r1, prod = r01 * r2;
      if (fabs(prod) > 1e-280 && fabs(prod) < 1e280) {
        const double ip = 1.0 / prod;
        q0 = (r1 * r2) * ip;
        q1 = (r0 * r2) * ip;
        q2 = r01 * ip;
      } else {
        q0 = 1.0 / r0;
        q1 = 1.0 / r1;
        q2 = 1.0 / r2;
      }
      {
        const double zr = z[0] * q0;
        F += zr;
        err += fabs(zr);
        dF += zr * q0;
      }
      {
        const double zr = z[1] * q1;
        F += zr;
        err += fabs(zr);
        dF += zr * q1;
      }
      {
        const double zr = z[2] * q2;
        F += zr;
        err += fabs(zr);
        dF += zr * q2;
      }
      if (fabs(F) <= kEps * (8.0 * err + fabs(x * dF))) break;
      if (F < 0.0)
        xl = x;
      else
        xh = x;
      if (fabs(2.0 * F) < fabs(dxold * dF)) {
        dxold = dx;
        dx = F / dF;
        x = x - dx;
        if ((x - xh) * (x - xl) >= 0.0) {
          dx = 0.5 * (xh - xl);
          x = xl + dx;
        }
      } else {
        dx = 0.5 * (xh - xl);
        x = xl + dx;
      }
      r0 = dd0 - x;
      r1 = dd1 - x;
      r2 = dd2 - x;
    }
    P[0][i] = r0;
    P[1][i] = r1;
    P[2][i] = r2;
    w[i] = x + delta;
  }
}

// Eigenvalues w and column eigenvectors Q of symmetric A (upper part read).
__device__ __forceinline__ void eig_sym3(const double A[3][3], double Q[3][3], double w[3]) {
  double Hh[3][3], P[3][3], e[2], d[3], z[3], c, s, t;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Q[i][j] = 0.0;
  eig_tridiag(A, Hh, w, e);
  // split when an off-diagonal element vanishes
  t = fabs(w[0]) + fabs(w[1]);
  if (fabs(e[0]) <= 8.0 * kEps * t) {
    eig_2x2(w[1], e[1], w[2], d[1], d[2], c, s);
    w[1] = d[1];
    w[2] = d[2];
    Q[0][0] = 1.0;
#pragma unroll
    for (int j = 1; j < 3; j++) {
      Q[j][1] = s * Hh[j][2] + c * Hh[j][1];
      Q[j][2] = c * Hh[j][2] - s * Hh[j][1];
    }
    return;
  }
  t = fabs(w[1]) + fabs(w[2]);
  if (fabs(e[1]) <= 8.0 * kEps * t) {
    eig_2x2(w[0], e[0], w[1], d[0], d[1], c, s);
    w[0] = d[0];
    w[1] = d[1];
    Q[0][0] = c;
    Q[0][1] = -s;
    Q[1][0] = Hh[1][1] * s;
    Q[1][1] = Hh[1][1] * c;
    Q[1][2] = Hh[1][2];
    Q[2][0] = Hh[2][1] * s;
    Q[2][1] = Hh[2][1] * c;
    Q[2][2] = Hh[2][2];
    return;
  }
  eig_2x2(w[1] - e[0], e[1], w[2], d[1], d[2], c, s);
  d[0] = w[0] - e[0];
  z[0] = e[0];
  z[1] = e[0] * sqd(c);
  z[2] = e[0] * sqd(s);
  if (d[0] < d[2])
    eig_secular<0, 2, 1>(d, z, w, P);
  else if (d[0] < d[1])
    eig_secular<2, 0, 1>(d, z, w, P);
  else
    eig_secular<2, 1, 0>(d, z, w, P);
  z[0] = sqrt(fabs(e[0]));
  z[1] = c * z[0];
  z[2] = -s * z[0];
  t = 8.0 * kEps * (fabs(d[0]) + fabs(d[1]) + fabs(d[2]));
  if (fabs(d[1] - d[0]) <= t) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (P[0][j] * P[1][j] <= 0.0) {
        P[0][j] = z[1];
        P[1][j] = -z[0];
        P[2][j] = 0.0;
      } else {
        P[0][j] = z[0] / P[0][j];
        P[1][j] = z[1] / P[1][j];
        P[2][j] = z[2] / P[2][j];
      }
    }
  } else if (fabs(d[2] - d[0]) <= t) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (P[0][j] * P[2][j] <= 0.0) {
        P[0][j] = z[2];
        P[1][j] = 0.0;
        P[2][j] = -z[0];
      } else {
        P[0][j] = z[0] / P[0][j];
        P[1][j] = z[1] / P[1][j];
        P[2][j] = z[2] / P[2][j];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      // first zero entry of the column makes it a unit vector
      if (P[0][j] == 0.0) {
        P[0][j] = 1.0;
        P[1][j] = 0.0;
        P[2][j] = 0.0;
      } else {
        P[0][j] = z[0] / P[0][j];
        if (P[1][j] == 0.0) {
          P[1][j] = 1.0;
          P[2][j] = 0.0;
          P[0][j] = 0.0;
        } else {
          P[1][j] = z[1] / P[1][j];
          if (P[2][j] == 0.0) {
            P[2][j] = 1.0;
            P[0][j] = 0.0;
            P[1][j] = 0.0;
          } else {
            P[2][j] = z[2] / P[2][j];
          }
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    t = sqd(P[0][j]) + sqd(P[1][j]) + sqd(P[2][j]);
    t = 1.0 / sqrt(t);
    P[0][j] *= t;
    P[1][j] *= t;
    P[2][j] *= t;
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    t = P[1][j];
    P[1][j] = c * t - s * P[2][j];
    P[2][j] = s * t + c * P[2][j];
  }
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      t = P[k][j];
#pragma unroll
      for (int i = 0; i < 3; i++) Q[i][j] += t * Hh[i][k];
    }
}

// Cyclic Jacobi eigen-decomposition (Golub & Van Loan, Matrix Computations,
// 4th ed., Algorithms 8.5.1-8.5.3).  Not part of the reference: taken only
// where eig_sym3 (like the reference's dsyevd3) returns non-finite or
// non-orthonormal eigenvectors, which happens for nearly repeated
// eigenvalue pairs (e.g. the equal lateral stretches of a uniaxial wave).
// Columns of Q are eigenvectors.
__device__ __forceinline__ void jacobi_eig3(const double A[3][3], double Q[3][3], double w[3]) {
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a[i][j] = A[i][j];
      Q[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 32; sweep++) {
    const double off = sqd(a[0][1]) + sqd(a[0][2]) + sqd(a[1][2]);
    const double nrm = sqd(a[0][0]) + sqd(a[1][1]) + sqd(a[2][2]) + 2.0 * off;
    if (off <= 1e-32 * nrm) break;
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int p = (r == 2) ? 1 : 0, q = (r == 0) ? 1 : 2;
      if (a[p][q] == 0.0) continue;
      const double tau = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq;
        a[k][q] = s * akp + c * akq;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk;
        a[q][k] = s * apk + c * aqk;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double vkp = Q[k][p], vkq = Q[k][q];
        Q[k][p] = c * vkp - s * vkq;
        Q[k][q] = s * vkp + c * vkq;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; i++) w[i] = a[i][i];
}

// F^T F of F (row-major 9), as Tensor2::buildFTF (Tensor2.C:1315-1327).
__device__ __forceinline__ void build_ftf(const double F[9], double C[3][3]) {
  C[0][0] = sqd(F[0]) + sqd(F[3]) + sqd(F[6]);
  C[0][1] = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  C[0][2] = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  C[1][1] = sqd(F[1]) + sqd(F[4]) + sqd(F[7]);
  C[1][2] = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  C[2][2] = sqd(F[2]) + sqd(F[5]) + sqd(F[8]);
  C[1][0] = C[0][1];
  C[2][0] = C[0][2];
  C[2][1] = C[1][2];
}

// Tensor2::polarExtractLnU (Tensor2.C:1237-1312): from the eigen pairs of
// F^T F, dE = 1/2 sum ln(lam_k) v_k v_k^T and R = F U^-1.
__device__ __forceinline__ void polar_extract(const double F[9], const double V[3][3], const double lam[3],
                                              double dE[6], double R[9]) {
  double D[3][6];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    D[k][0] = sqd(V[0][k]);
    D[k][1] = V[0][k] * V[1][k];
    D[k][2] = V[0][k] * V[2][k];
    D[k][3] = sqd(V[1][k]);
    D[k][4] = V[1][k] * V[2][k];
    D[k][5] = sqd(V[2][k]);
  }
  const double g0 = sqrt(lam[0]), g1 = sqrt(lam[1]), g2 = sqrt(lam[2]);
  double U[6];
#pragma unroll
  for (int c = 0; c < 6; c++) U[c] = g0 * D[0][c] + g1 * D[1][c] + g2 * D[2][c];
  const double t1 = U[3] * U[5], t2 = U[2] * U[4], t4 = U[4] * U[4], t5 = U[1] * U[5], t6 = U[2] * U[3];
  const double det = U[0] * t1 + 2.0 * U[1] * t2 - U[0] * t4 - U[1] * t5 - U[2] * t6;
  const double id = 1.0 / det;
  double Ui[6];
  Ui[0] = (t1 - t4) * id;
  Ui[1] = (t2 - t5) * id;
  Ui[2] = (U[1] * U[4] - t6) * id;
  Ui[3] = (U[0] * U[5] - U[2] * U[2]) * id;
  Ui[4] = (U[2] * U[1] - U[0] * U[4]) * id;
  Ui[5] = (U[0] * U[3] - U[1] * U[1]) * id;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double a0 = F[3 * r], a1 = F[3 * r + 1], a2 = F[3 * r + 2];
    R[3 * r + 0] = a0 * Ui[0] + a1 * Ui[1] + a2 * Ui[2];
    R[3 * r + 1] = a0 * Ui[1] + a1 * Ui[3] + a2 * Ui[4];
    R[3 * r + 2] = a0 * Ui[2] + a1 * Ui[4] + a2 * Ui[5];
  }
  const double l0 = log(lam[0]) * 0.5, l1 = log(lam[1]) * 0.5, l2 = log(lam[2]) * 0.5;
#pragma unroll
  for (int c = 0; c < 6; c++) dE[c] = l0 * D[0][c] + l1 * D[1][c] + l2 * D[2][c];
}

// Tensor2::polarCuppenLnU (Tensor2.C:1442-1457) on F (row-major 9):
// dE = ln U (sym 6), R = F U^-1 (row-major 9).  Where the Kopp solver
// returns non-finite or non-orthonormal eigenvectors the Jacobi
// decomposition replaces it.
__device__ __forceinline__ void polar_lnU_fp64(const double F[9], double dE[6], double R[9]) {
  double C[3][3];
  build_ftf(F, C);
  double V[3][3], lam[3];
  eig_sym3(C, V, lam);
  // usable: finite, eigenvector columns orthonormal to 64 eps (the Cuppen
  // step loses orthogonality for nearly repeated eigenvalues, which makes
  // U and R wrong even for tiny strains)
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    ok = ok && isfinite(lam[i]);
#pragma unroll
    for (int j = i; j < 3; j++) {
      const double d = V[0][i] * V[0][j] + V[1][i] * V[1][j] + V[2][i] * V[2][j] - (i == j ? 1.0 : 0.0);
      ok = ok && (fabs(d) <= 64.0 * kEps);
    }
  }
  if (!ok) jacobi_eig3(C, V, lam);
  polar_extract(F, V, lam, dE, R);
}

// FP64 polar decomposition through the Jacobi solver alone (no reference
// arithmetic to follow): the large-increment branch of the mixed path.
__device__ __forceinline__ void polar_lnU_jacobi(const double F[9], double dE[6], double R[9]) {
  double C[3][3];
  build_ftf(F, C);
  double V[3][3], lam[3];
  jacobi_eig3(C, V, lam);
  polar_extract(F, V, lam, dE, R);
}

// R . S . R^T for symmetric S (SymTensor2::dotRxRT, SymTensor2.C:512-525).
template <typename T>
__device__ __forceinline__ void rotate_sym(const T r[9], T S[6]) {
  T o[6];
  o[0] = r[0] * r[0] * S[0] + 2 * r[0] * (r[1] * S[1] + r[2] * S[2]) + r[1] * r[1] * S[3] + 2 * r[1] * r[2] * S[4] +
         r[2] * r[2] * S[5];
  o[1] = r[0] * (r[3] * S[0] + r[4] * S[1] + r[5] * S[2]) + r[1] * (r[3] * S[1] + r[4] * S[3] + r[5] * S[4]) +
         r[2] * (r[3] * S[2] + r[4] * S[4] + r[5] * S[5]);
  o[2] = r[0] * (r[6] * S[0] + r[7] * S[1] + r[8] * S[2]) + r[1] * (r[6] * S[1] + r[7] * S[3] + r[8] * S[4]) +
         r[2] * (r[6] * S[2] + r[7] * S[4] + r[8] * S[5]);
  o[3] = r[3] * r[3] * S[0] + 2 * r[3] * (r[4] * S[1] + r[5] * S[2]) + r[4] * r[4] * S[3] + 2 * r[4] * r[5] * S[4] +
         r[5] * r[5] * S[5];
  o[4] = r[3] * (r[6] * S[0] + r[7] * S[1] + r[8] * S[2]) + r[4] * (r[6] * S[1] + r[7] * S[3] + r[8] * S[4]) +
         r[5] * (r[6] * S[2] + r[7] * S[4] + r[8] * S[5]);
  o[5] = r[6] * r[6] * S[0] + 2 * r[6] * (r[7] * S[1] + r[8] * S[2]) + r[7] * r[7] * S[3] + 2 * r[7] * r[8] * S[4] +
         r[8] * r[8] * S[5];
#pragma unroll
  for (int c = 0; c < 6; c++) S[c] = o[c];
}

}  // namespace dnlg
