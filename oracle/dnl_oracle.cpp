/*
 * dnl_oracle.cpp -- CPU FP64 oracle for the explicit solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The engine (libdnlg.so) never links or calls it.
 *
 * It restates, in plain serial C++ (optional OpenMP over independent
 * elements), the time-step loop of DynELA's Explicit::solve and everything it
 * calls for the ElQua4NAx / ElHex8N3D elements with the J2 radial return.
 * Loop order and arithmetic order follow the reference function by
 * function; each function cites the reference lines it restates.  It takes
 * the same descriptors as the engine's C ABI (include/dnlg.h), so a test
 * can feed both the same model.
 *
 * Pinning (tests/test_oracle_*.py):
 *  - kopp_eig3 is checked bit-for-bit against the reference's own dsyevd3
 *    compiled from Sources/dnlMaths/Eigen3x3.C into oracle/_ref;
 *  - shape tables vs the closed-form shape functions, partition of unity;
 *  - polar/log strain vs scipy polar/logm and invariants;
 *  - Johnson-Cook yield and derivative vs closed forms and the reference's
 *    finite-difference check; the radial return vs a brute-force bisection;
 *  - Chung-Hulbert constants vs closed forms; lumped mass vs rho*V;
 *    Courant lengths vs box/rectangle closed forms; free flight of a body
 *    with f = 0; elastic simple shear vs the Green-Naghdi closed form;
 *  - the ANN2Sig network law: its yield stress vs an independent forward
 *    pass on the weights as stored in the law's .npz file, its slope vs
 *    central differences of the yield stress, and a constant network's
 *    solve vs the same solve with a constant power law.
 *  Parity unpinned: the CBLAS dgemm summation order behind dN/dx
 *  (Matrix.C:815) -- a plain k-ordered sum is used.
 *
 * Deliberate extension: where kopp_eig3 (like the reference's dsyevd3)
 * returns non-finite eigenvectors, or eigenvectors that are not orthonormal
 * to 64 eps -- nearly repeated eigenvalue pairs -- polar_lnU falls back to a
 * cyclic Jacobi decomposition instead of propagating NaN or a wrong U and R;
 * the engine does the same.
 */
#include "../include/dnlg.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* constants of Sources/dnlKernel/Macros.h:37-56 */
const double kPI = 3.14159265358979324;
const double k2PI = 6.28318530717958648;
const double kSqrt3 = 1.73205080756887729;
const double kSqrt23 = 0.816496580927726034;
const double kSqrt32 = 1.22474487139158894;
const double kEps = 2.2204460492503131e-16;

inline double sq(double a) { return a * a; }

/* ---------------------------------------------------------------------------
 * Symmetric 3x3 eigen-decomposition: Cuppen divide & conquer after a
 * Householder tridiagonalisation, the published method of J. Kopp,
 * "Efficient numerical diagonalization of hermitian 3x3 matrices",
 * arXiv:physics/0610206, as used by the reference in
 * Sources/dnlMaths/Eigen3x3.C:14-88 (tridiagonalisation), :194-383 (secular
 * equation), :385-453 (2x2 block), :454-634 (driver).  The arithmetic order
 * is kept identical so results agree bit-for-bit with the reference build.
 * ------------------------------------------------------------------------- */

/* A = Q T Q^T with T tridiagonal (diag d, off-diag e); reads upper part. */
void kopp_tridiag(const double A[3][3], double Q[3][3], double d[3], double e[2]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Q[i][j] = (i == j) ? 1.0 : 0.0;
  const double h = sq(A[0][1]) + sq(A[0][2]);
  const double g = (A[0][1] > 0) ? -std::sqrt(h) : std::sqrt(h);
  e[0] = g;
  double f = g * A[0][1];
  double u[3], q[3];
  u[1] = A[0][1] - g;
  u[2] = A[0][2];
  double omega = h - f;
  if (omega > 0.0) {
    omega = 1.0 / omega;
    double K = 0.0;
    for (int i = 1; i < 3; i++) {
      f = A[1][i] * u[1] + A[i][2] * u[2];
      q[i] = omega * f;
      K += u[i] * f;
    }
    K *= 0.5 * sq(omega);
    for (int i = 1; i < 3; i++) q[i] = q[i] - K * u[i];
    d[0] = A[0][0];
    d[1] = A[1][1] - 2.0 * q[1] * u[1];
    d[2] = A[2][2] - 2.0 * q[2] * u[2];
    for (int j = 1; j < 3; j++) {
      f = omega * u[j];
      for (int i = 1; i < 3; i++) Q[i][j] = Q[i][j] - f * u[i];
    }
    e[1] = A[1][2] - q[1] * u[2] - u[1] * q[2];
  } else {
    for (int i = 0; i < 3; i++) d[i] = A[i][i];
    e[1] = A[1][2];
  }
}

/* eigen-system of [[a,b],[b,c]]: rt1 >= rt2, vector (cs, sn) of rt1 */
void kopp_eig2(double a, double b, double c, double *rt1, double *rt2, double *cs, double *sn) {
  const double sm = a + c;
  const double df = a - c;
  const double rt = std::sqrt(sq(df) + 4.0 * b * b);
  double t;
  if (sm > 0.0) {
    *rt1 = 0.5 * (sm + rt);
    t = 1.0 / (*rt1);
    *rt2 = (a * t) * c - (b * t) * b;
  } else if (sm < 0.0) {
    *rt2 = 0.5 * (sm - rt);
    t = 1.0 / (*rt2);
    *rt1 = (a * t) * c - (b * t) * b;
  } else {
    *rt1 = 0.5 * rt;
    *rt2 = -0.5 * rt;
  }
  *cs = (df > 0.0) ? df + rt : df - rt;
  if (std::fabs(*cs) > 2.0 * std::fabs(b)) {
    t = -2.0 * b / *cs;
    *sn = 1.0 / std::sqrt(1.0 + sq(t));
    *cs = t * (*sn);
  } else if (std::fabs(b) == 0.0) {
    *cs = 1.0;
    *sn = 0.0;
  } else {
    t = -0.5 * (*cs) / b;
    *cs = 1.0 / std::sqrt(1.0 + sq(t));
    *sn = t * (*cs);
  }
  if (df > 0.0) {
    t = *cs;
    *cs = -(*sn);
    *sn = t;
  }
}

/* Roots w of 1 + sum z_i/(d_i - w) = 0 with d[i0] <= d[i1] <= d[i2];
 * P[j][i] = d_j - w_i. */
void kopp_secular(const double d[3], const double z[3], double w[3], double P[3][3], int i0, int i1,
                  int i2) {
  double a[4];
  if (z[0] > 0) {
    a[0] = d[i0];
    a[1] = d[i1];
    a[2] = d[i2];
    a[3] = std::fabs(d[0] + 3.0 * z[0]) + std::fabs(d[1] + 3.0 * z[1]) + std::fabs(d[2] + 3.0 * z[2]);
  } else {
    a[0] = -std::fabs(d[0] + 3.0 * z[0]) - std::fabs(d[1] + 3.0 * z[1]) - std::fabs(d[2] + 3.0 * z[2]);
    a[1] = d[i0];
    a[2] = d[i1];
    a[3] = d[i2];
  }
  /* closed-form cubic roots as starting values */
  double t0 = d[1] * d[2], t1 = d[0] * d[2], t2 = d[0] * d[1];
  const double cg = t0 * d[0] + (z[0] * t0 + z[1] * t1 + z[2] * t2);
  const double cb = (z[0] * (d[1] + d[2]) + z[1] * (d[0] + d[2]) + z[2] * (d[0] + d[1])) + (t0 + t1 + t2);
  const double ca = (z[0] + z[1] + z[2]) + (d[0] + d[1] + d[2]);
  const double p = sq(ca) - 3.0 * cb;
  const double q = ca * (p - (3.0 / 2.0) * cb) + (27.0 / 2.0) * cg;
  const double sqrt_p = std::sqrt(std::fabs(p));
  double phi = 27.0 * (0.25 * sq(cb) * (p - cb) - cg * (q - 27.0 / 4.0 * cg));
  phi = (1.0 / 3.0) * std::atan2(std::sqrt(std::fabs(phi)), q);
  const double c = sqrt_p * std::cos(phi);
  const double s = (1.0 / kSqrt3) * sqrt_p * std::fabs(std::sin(phi));
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
  /* safeguarded Newton / bisection refinement of each root */
  for (int i = 0; i < 3; i++) {
    double xl = a[i], xh = a[i + 1];
    double dx = 0.5 * (xh - xl), dxold = dx;
    double dd[3], x, delta;
    if (dx == 0.0) {
      w[i] = xl;
      for (int j = 0; j < 3; j++) P[j][i] = d[j] - xl;
      continue;
    }
    if (x0[i] >= xh) {
      delta = xh;
      x = -dx;
      for (int j = 0; j < 3; j++) {
        dd[j] = d[j] - delta;
        P[j][i] = dd[j] - x;
      }
    } else if (x0[i] <= xl) {
      delta = xl;
      x = dx;
      for (int j = 0; j < 3; j++) {
        dd[j] = d[j] - delta;
        P[j][i] = dd[j] - x;
      }
    } else {
      delta = x0[i];
      x = 0.0;
      for (int j = 0; j < 3; j++) P[j][i] = dd[j] = d[j] - delta;
    }
    xl -= delta;
    xh -= delta;
    if (z[0] < 0.0) {
      const double tmp = xh;
      xh = xl;
      xl = tmp;
    }
    for (int it = 0; it < 500; it++) {
      double F = 1.0, dF = 0.0, err = 1.0;
      for (int j = 0; j < 3; j++) {
        const double r = 1.0 / P[j][i];
        const double zr = z[j] * r;
        const double zrr = zr * r;
        F += zr;
        err += std::fabs(zr);
        dF += zrr;
      }
      if (std::fabs(F) <= kEps * (8.0 * err + std::fabs(x * dF))) break;
      if (F < 0.0)
        xl = x;
      else
        xh = x;
      if (std::fabs(2.0 * F) < std::fabs(dxold * dF)) {
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
      for (int j = 0; j < 3; j++) P[j][i] = dd[j] - x;
    }
    w[i] = x + delta;
  }
}

/* eigenvalues w and column eigenvectors Q of the symmetric A */
void kopp_eig3(const double A[3][3], double Q[3][3], double w[3]) {
  double Hh[3][3], P[3][3], e[2], d[3], z[3], c, s, t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Q[i][j] = 0.0;
  kopp_tridiag(A, Hh, w, e);
  /* split when an off-diagonal vanishes */
  for (int i = 0; i < 2; i++) {
    t = std::fabs(w[i]) + std::fabs(w[i + 1]);
    if (std::fabs(e[i]) <= 8.0 * kEps * t) {
      if (i == 0) {
        kopp_eig2(w[1], e[1], w[2], &d[1], &d[2], &c, &s);
        w[1] = d[1];
        w[2] = d[2];
        Q[0][0] = 1.0;
        for (int j = 1; j < 3; j++) {
          Q[j][1] = s * Hh[j][2] + c * Hh[j][1];
          Q[j][2] = c * Hh[j][2] - s * Hh[j][1];
        }
      } else {
        kopp_eig2(w[0], e[0], w[1], &d[0], &d[1], &c, &s);
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
      }
      return;
    }
  }
  /* divide: 2x2 block of T - e0 v v^T */
  kopp_eig2(w[1] - e[0], e[1], w[2], &d[1], &d[2], &c, &s);
  d[0] = w[0] - e[0];
  /* conquer: rank-one update, secular equation */
  z[0] = e[0];
  z[1] = e[0] * sq(c);
  z[2] = e[0] * sq(s);
  if (d[0] < d[2])
    kopp_secular(d, z, w, P, 0, 2, 1);
  else if (d[0] < d[1])
    kopp_secular(d, z, w, P, 2, 0, 1);
  else
    kopp_secular(d, z, w, P, 2, 1, 0);
  z[0] = std::sqrt(std::fabs(e[0]));
  z[1] = c * z[0];
  z[2] = -s * z[0];
  t = 8.0 * kEps * (std::fabs(d[0]) + std::fabs(d[1]) + std::fabs(d[2]));
  if (std::fabs(d[1] - d[0]) <= t) {
    for (int j = 0; j < 3; j++) {
      if (P[0][j] * P[1][j] <= 0.0) {
        P[0][j] = z[1];
        P[1][j] = -z[0];
        P[2][j] = 0.0;
      } else
        for (int i = 0; i < 3; i++) P[i][j] = z[i] / P[i][j];
    }
  } else if (std::fabs(d[2] - d[0]) <= t) {
    for (int j = 0; j < 3; j++) {
      if (P[0][j] * P[2][j] <= 0.0) {
        P[0][j] = z[2];
        P[1][j] = 0.0;
        P[2][j] = -z[0];
      } else
        for (int i = 0; i < 3; i++) P[i][j] = z[i] / P[i][j];
    }
  } else {
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        if (P[i][j] == 0.0) {
          P[i][j] = 1.0;
          P[(i + 1) % 3][j] = 0.0;
          P[(i + 2) % 3][j] = 0.0;
          break;
        } else
          P[i][j] = z[i] / P[i][j];
      }
  }
  for (int j = 0; j < 3; j++) {
    t = sq(P[0][j]) + sq(P[1][j]) + sq(P[2][j]);
    t = 1.0 / std::sqrt(t);
    for (int i = 0; i < 3; i++) P[i][j] *= t;
  }
  for (int j = 0; j < 3; j++) {
    t = P[1][j];
    P[1][j] = c * t - s * P[2][j];
    P[2][j] = s * t + c * P[2][j];
  }
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) {
      t = P[k][j];
      for (int i = 0; i < 3; i++) Q[i][j] += t * Hh[i][k];
    }
}

/* ---------------------------------------------------------------------------
 * Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van
 * Loan, Matrix Computations, 4th ed., Algorithms 8.5.1-8.5.3): rotations in
 * the (0,1), (0,2), (1,2) planes until the off-diagonal mass is below
 * 1e-32 of the Frobenius mass.  Not part of the reference: used only where
 * kopp_eig3 (like the reference's dsyevd3) returns non-finite or
 * non-orthonormal eigenvectors, which happens for nearly repeated
 * eigenvalue pairs (e.g. the equal lateral stretches of a uniaxial wave).
 * Q holds the eigenvectors as columns.
 * ------------------------------------------------------------------------- */
void jacobi_eig3(const double A[3][3], double Q[3][3], double w[3]) {
  double a[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      a[i][j] = A[i][j];
      Q[i][j] = (i == j) ? 1.0 : 0.0;
    }
  const int P[3] = {0, 0, 1}, Qi[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 32; sweep++) {
    const double off = sq(a[0][1]) + sq(a[0][2]) + sq(a[1][2]);
    const double nrm = sq(a[0][0]) + sq(a[1][1]) + sq(a[2][2]) + 2.0 * off;
    if (off <= 1e-32 * nrm) break;
    for (int r = 0; r < 3; r++) {
      const int p = P[r], q = Qi[r];
      if (a[p][q] == 0.0) continue;
      const double tau = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
      const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
      for (int k = 0; k < 3; k++) {
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq;
        a[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk;
        a[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {
        const double vkp = Q[k][p], vkq = Q[k][q];
        Q[k][p] = c * vkp - s * vkq;
        Q[k][q] = s * vkp + c * vkq;
      }
    }
  }
  for (int i = 0; i < 3; i++) w[i] = a[i][i];
}

/* Eigen pairs usable for the polar decomposition: finite, and eigenvector
 * columns orthonormal to 64 eps (the Cuppen step loses orthogonality for
 * nearly repeated eigenvalues, which turns U = sum g_k v_k v_k^T and
 * R = F U^-1 into wrong tensors even for tiny strains). */
bool usable_eigen(const double Q[3][3], const double w[3]) {
  bool ok = true;
  for (int i = 0; i < 3; i++) {
    ok = ok && std::isfinite(w[i]);
    for (int j = 0; j < 3; j++) ok = ok && std::isfinite(Q[i][j]);
  }
  if (!ok) return false;
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++) {
      const double d = Q[0][i] * Q[0][j] + Q[1][i] * Q[1][j] + Q[2][i] * Q[2][j] - (i == j ? 1.0 : 0.0);
      if (!(std::fabs(d) <= 64.0 * kEps)) return false;
    }
  return true;
}

/* ---------------------------------------------------------------------------
 * Polar decomposition F = R U and logarithmic strain increment ln U:
 * Tensor2::buildFTF (Tensor2.C:1315-1327), Tensor2::polarExtractLnU
 * (Tensor2.C:1237-1312), Tensor2::polarCuppenLnU (Tensor2.C:1442-1457).
 * F row-major [9]; dE symmetric [6]; R row-major [9].
 * ------------------------------------------------------------------------- */
void polar_lnU(const double F[9], double dE[6], double R[9]) {
  double C[3][3];
  C[0][0] = sq(F[0]) + sq(F[3]) + sq(F[6]);
  C[0][1] = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  C[0][2] = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  C[1][0] = C[0][1];
  C[1][1] = sq(F[1]) + sq(F[4]) + sq(F[7]);
  C[1][2] = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  C[2][0] = C[0][2];
  C[2][1] = C[1][2];
  C[2][2] = sq(F[2]) + sq(F[5]) + sq(F[8]);
  double V[3][3], lam[3];
  kopp_eig3(C, V, lam);
  if (!usable_eigen(V, lam)) jacobi_eig3(C, V, lam);
  /* dyads v_k (x) v_k, symmetric storage */
  double D[3][6];
  for (int k = 0; k < 3; k++) {
    D[k][0] = sq(V[0][k]);
    D[k][1] = V[0][k] * V[1][k];
    D[k][2] = V[0][k] * V[2][k];
    D[k][3] = sq(V[1][k]);
    D[k][4] = V[1][k] * V[2][k];
    D[k][5] = sq(V[2][k]);
  }
  double g[3] = {std::sqrt(lam[0]), std::sqrt(lam[1]), std::sqrt(lam[2])};
  double U[6];
  for (int c = 0; c < 6; c++) U[c] = g[0] * D[0][c] + g[1] * D[1][c] + g[2] * D[2][c];
  /* closed-form inverse of the symmetric U */
  const double t1 = U[3] * U[5];
  const double t2 = U[2] * U[4];
  const double t4 = U[4] * U[4];
  const double t5 = U[1] * U[5];
  const double t6 = U[2] * U[3];
  const double det = U[0] * t1 + 2.0 * U[1] * t2 - U[0] * t4 - U[1] * t5 - U[2] * t6;
  double Ui[6];
  Ui[0] = (t1 - t4) / det;
  Ui[1] = (t2 - t5) / det;
  Ui[2] = (U[1] * U[4] - t6) / det;
  Ui[3] = (U[0] * U[5] - U[2] * U[2]) / det;
  Ui[4] = (U[2] * U[1] - U[0] * U[4]) / det;
  Ui[5] = (U[0] * U[3] - U[1] * U[1]) / det;
  for (int r = 0; r < 3; r++) {
    const double a0 = F[3 * r], a1 = F[3 * r + 1], a2 = F[3 * r + 2];
    R[3 * r + 0] = a0 * Ui[0] + a1 * Ui[1] + a2 * Ui[2];
    R[3 * r + 1] = a0 * Ui[1] + a1 * Ui[3] + a2 * Ui[4];
    R[3 * r + 2] = a0 * Ui[2] + a1 * Ui[4] + a2 * Ui[5];
  }
  g[0] = std::log(lam[0]) / 2;
  g[1] = std::log(lam[1]) / 2;
  g[2] = std::log(lam[2]) / 2;
  for (int c = 0; c < 6; c++) dE[c] = g[0] * D[0][c] + g[1] * D[1][c] + g[2] * D[2][c];
}

/* R . S . R^T for symmetric S (SymTensor2::dotRxRT, SymTensor2.C:512-525) */
void rotate_sym(const double R[9], double S[6]) {
  const double *r = R;
  double o[6];
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
  for (int c = 0; c < 6; c++) S[c] = o[c];
}

/* SymTensor2 helpers: thirdTrace (SymTensor2.h:327-331), deviator
 * (SymTensor2.C:688-699), norm (SymTensor2.h:362-367), doubleDot
 * (SymTensor2.C:640-649). */
inline double third_trace(const double S[6]) { return (S[0] + S[3] + S[5]) / 3.0; }
inline void deviator(const double S[6], double out[6]) {
  const double p = third_trace(S);
  for (int c = 0; c < 6; c++) out[c] = S[c];
  out[0] -= p;
  out[3] -= p;
  out[5] -= p;
}
inline double norm_sym(const double S[6]) {
  return std::sqrt(sq(S[0]) + sq(S[3]) + sq(S[5]) + 2.0 * (sq(S[1]) + sq(S[2]) + sq(S[4])));
}
inline double ddot_sym(const double A[6], const double B[6]) {
  return (A[0] * B[0] + A[3] * B[3] + A[5] * B[5] + 2.0 * (A[1] * B[1] + A[2] * B[2] + A[4] * B[4]));
}
/* index of (i,j) in [xx,xy,xz,yy,yz,zz] */
inline int sym_idx(int i, int j) {
  static const int m[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  return m[i][j];
}

/* ---------------------------------------------------------------------------
 * Hardening laws.
 * ------------------------------------------------------------------------- */
struct Law {
  int kind;
  double p[8];
  double eta, rho, cp; /* Taylor-Quinney, density, heat capacity */
  std::vector<double> tx, ty; /* DNLG_LAW_TABULAR points */
  std::vector<double> ann;    /* DNLG_LAW_ANN2SIG packed network */
};

Law law_of(const dnlg_material *m) {
  Law L{m->law, {0}, m->taylor_quinney, m->density, m->heat_capacity};
  for (int k = 0; k < 8; k++) L.p[k] = m->p[k];
  if (m->law == DNLG_LAW_TABULAR && m->table)
    for (int64_t k = 0; k < m->n_table; k++) {
      L.tx.push_back(m->table[2 * k]);
      L.ty.push_back(m->table[2 * k + 1]);
    }
  if (m->law == DNLG_LAW_ANN2SIG && m->table) L.ann.assign(m->table, m->table + m->n_table);
  return L;
}

/* ANN2SigLaw::getYieldStress / getDerivateYieldStress (ANN2SigLaw.C:173-245),
 * layer by layer as the reference's vector expressions: inputs scaled by
 * minEntries / rangeEntries (rate: log(depsp / logBase), 0 at or below
 * logBase, where depsp becomes logBase), a = exp(-(w1 x + b1)),
 * s1 = 1 / (1 + a), c = exp(-(w2 s1 + b2)), y = w3 / (1 + c) + b3,
 * yield = range3 y + min3; the slope back-propagates
 * d = w3 c / (1 + c)^2, f = (w2^T d) a / (1 + a)^2, yd = w1^T f. */
double ann(const Law &L, double epsp, double depsp, double T, double dtime, bool slope) {
  if (L.ann.size() < 2) return std::nan("");
  const double *t = L.ann.data();
  const int h1 = (int)t[0], h2 = (int)t[1];
  if (h1 < 1 || h2 < 1 || (int64_t)L.ann.size() != DNLG_ANN_TABLE_SIZE((int64_t)h1, (int64_t)h2))
    return std::nan("");
  const double *w1 = t + 2, *b1 = w1 + 3 * h1, *w2 = b1 + h1, *b2 = w2 + h2 * h1, *w3 = b2 + h2;
  const double b3 = w3[h2], logBase = w3[h2 + 1];
  const double *mn = w3 + h2 + 2, *mx = mn + 4;
  double range[4];
  for (int k = 0; k < 4; k++) range[k] = mx[k] - mn[k];
  double in[3];
  in[0] = (epsp - mn[0]) / range[0];
  if (depsp > logBase) {
    in[1] = (std::log(depsp / logBase) - mn[1]) / range[1];
  } else {
    in[1] = 0.0;
    depsp = logBase;
  }
  in[2] = (T - mn[2]) / range[2];
  std::vector<double> a(h1), s1(h1), c(h2), s2(h2);
  for (int i = 0; i < h1; i++) {
    double z = 0.0;
    for (int k = 0; k < 3; k++) z += w1[3 * i + k] * in[k];
    a[i] = std::exp(-(z + b1[i]));
    s1[i] = 1.0 / (a[i] + 1.0);
  }
  double y = 0.0;
  for (int j = 0; j < h2; j++) {
    double z = 0.0;
    for (int i = 0; i < h1; i++) z += w2[j * h1 + i] * s1[i];
    c[j] = std::exp(-(z + b2[j]));
    s2[j] = 1.0 / (c[j] + 1.0);
    y += w3[j] * s2[j];
  }
  y += b3;
  const double yield = range[3] * y + mn[3];
  if (!slope) return yield;
  std::vector<double> d(h2), f(h1);
  for (int j = 0; j < h2; j++) d[j] = 1.0 / ((c[j] + 1.0) * (c[j] + 1.0)) * c[j] * w3[j];
  for (int i = 0; i < h1; i++) {
    double g = 0.0;
    for (int j = 0; j < h2; j++) g += w2[j * h1 + i] * d[j];
    f[i] = g * (1.0 / ((a[i] + 1.0) * (a[i] + 1.0)) * a[i]);
  }
  double yd[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < h1; i++) yd[k] += w1[3 * i + k] * f[i];
  const double dep = range[3] / range[0] * yd[0];
  const double drate = range[3] / range[1] * yd[1] / depsp;
  const double dT = range[3] / range[2] * yd[2];
  return dep + drate / dtime + L.eta / (L.rho * L.cp) * yield * dT;
}

/* DiscreteFunction::getValue / getSlope (DiscreteFunction.C:188-216,
 * 284-308) of a TabularLaw; NaN outside the table, where the reference
 * stops with a fatal error. */
double tabular(const Law &L, double x, bool slope) {
  const size_t n = L.tx.size();
  if (n < 2 || x < L.tx[0] || x > L.tx[n - 1]) return std::nan("");
  for (size_t i = 1; i < n; i++)
    if (L.tx[i] >= x) {
      if (slope) return (L.ty[i] - L.ty[i - 1]) / (L.tx[i] - L.tx[i - 1]);
      if (L.tx[i] == x) return L.ty[i];
      return ((x - L.tx[i - 1]) / (L.tx[i] - L.tx[i - 1])) * (L.ty[i] - L.ty[i - 1]) + L.ty[i - 1];
    }
  return 0.0;
}

/* JohnsonCookLaw::getYieldStress (JohnsonCookLaw.C:146-170) and
 * ElastoplasticLaw::getYieldStress (ElastoplasticLaw.C:110-114). */
double law_yield(const Law &L, double epsp, double depsp, double T) {
  if (L.kind == DNLG_LAW_TABULAR) return tabular(L, epsp, false);
  if (L.kind == DNLG_LAW_ANN2SIG) return ann(L, epsp, depsp, T, 0.0, false); /* dtime: slope only */
  if (L.kind == DNLG_LAW_JOHNSON_COOK) {
    const double A = L.p[0], B = L.p[1], C = L.p[2], n = L.p[3], m = L.p[4], d0 = L.p[5], Tm = L.p[6],
                 T0 = L.p[7];
    const double hard = A + B * std::pow(epsp, n);
    double visc;
    if (depsp > d0)
      visc = 1.0 + C * std::log(depsp / d0);
    else
      visc = 1.0;
    double temp = 1.0;
    if (T > T0) {
      if (T < Tm)
        temp = 1.0 - std::pow(((T - T0) / (Tm - T0)), m);
      else
        temp = 0.0;
    }
    return hard * visc * temp;
  }
  if (L.kind == DNLG_LAW_POWER) return (L.p[0] + L.p[1] * std::pow(epsp, L.p[2]));
  /* Voce + linear: no reference counterpart; closed form */
  return L.p[0] + L.p[1] * (1.0 - std::exp(-L.p[2] * epsp)) + L.p[3] * epsp;
}

/* JohnsonCookLaw::getDerivateYieldStress (JohnsonCookLaw.C:298-339) and
 * ElastoplasticLaw::getDerivateYieldStress (ElastoplasticLaw.C:117-121). */
double law_dyield(const Law &L, double epsp, double depsp, double T, double dtime) {
  if (L.kind == DNLG_LAW_TABULAR) return tabular(L, epsp, true);
  if (L.kind == DNLG_LAW_ANN2SIG) return ann(L, epsp, depsp, T, dtime, true);
  if (L.kind == DNLG_LAW_JOHNSON_COOK) {
    const double A = L.p[0], B = L.p[1], C = L.p[2], n = L.p[3], m = L.p[4], d0 = L.p[5], Tm = L.p[6],
                 T0 = L.p[7];
    double tp1 = 0.0;
    const double hard0 = A + B * std::pow(epsp, n);
    double visc0 = 1.0;
    if (depsp > d0) visc0 = 1.0 + C * std::log(depsp / d0);
    double temp0 = 1.0;
    if (T > T0) {
      if (T < Tm) {
        tp1 = std::pow(((T - T0) / (Tm - T0)), m);
        temp0 = 1.0 - tp1;
      } else
        temp0 = 0.0;
    }
    double h = n * B * std::pow(epsp, (n - 1.0)) * visc0 * temp0;
    if (depsp > d0) h += hard0 * C * temp0 / (depsp * dtime);
    if ((T > T0) && (T < Tm))
      h -= m * L.eta * hard0 * hard0 * visc0 * visc0 * temp0 * tp1 / (L.rho * L.cp * (T - T0));
    return h;
  }
  if (L.kind == DNLG_LAW_POWER) return (L.p[2] * L.p[1] * std::pow(epsp, (L.p[2] - 1.0)));
  return L.p[1] * L.p[2] * std::exp(-L.p[2] * epsp) + L.p[3];
}

/* ---------------------------------------------------------------------------
 * Element tables.  Values at the Gauss points are the closed forms the
 * reference tabulates (ElQua4NAx.C:63-110, ElHex8N3D.C:93-231): each shape
 * function value / derivative is a product of (1 +- 1/sqrt3) factors and is
 * selected here from the number of factors whose signs agree.
 * ------------------------------------------------------------------------- */
struct Tables {
  int type, dim, nn, nip;
  bool axi; /* ElQua4NAx (Element::Axisymetric); ElQua4N2D is plane strain */
  double w[8];
  double N[8][8];      /* [ip][node] */
  double dNdr[8][8][3]; /* [ip][node][local dir] */
};

Tables make_tables(int type) {
  Tables T{};
  T.type = type;
  T.axi = type == DNLG_EL_QUA4N_AX;
  if (type == DNLG_EL_QUA4N_AX || type == DNLG_EL_QUA4N_2D) {
    T.dim = 2;
    T.nn = 4;
    T.nip = 4;
    const int ns[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    const int ps[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    const double Nv[3] = {(2.0 - std::sqrt(3.0)) / 6, 1.0 / 6, (2.0 + std::sqrt(3.0)) / 6};
    const double Dv[2] = {(3.0 - std::sqrt(3.0)) / 12, (3.0 + std::sqrt(3.0)) / 12};
    for (int p = 0; p < 4; p++) {
      T.w[p] = 1.0;
      for (int a = 0; a < 4; a++) {
        int k = 0;
        for (int c = 0; c < 2; c++) k += (ns[a][c] == ps[p][c]);
        T.N[p][a] = Nv[k];
        for (int c = 0; c < 2; c++) {
          const int o = 1 - c;
          const double v = Dv[ns[a][o] == ps[p][o] ? 1 : 0];
          T.dNdr[p][a][c] = ns[a][c] > 0 ? v : -v;
        }
        T.dNdr[p][a][2] = 0.0;
      }
    }
  } else if (type == DNLG_EL_TET4N_3D) {
    /* ElTet4N3D.C:25-80: one point at the centroid, weight 1/6, linear
     * shape functions 1 - xi - eta - zeta, xi, eta, zeta */
    T.dim = 3;
    T.nn = 4;
    T.nip = 1;
    T.w[0] = 1.0 / 6;
    const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 4; a++) {
      T.N[0][a] = 1.0 / 4;
      for (int c = 0; c < 3; c++) T.dNdr[0][a][c] = d[a][c];
    }
  } else {
    T.dim = 3;
    T.nn = 8;
    T.nip = 8;
    const int ns[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                          {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    const double Nv[4] = {(9.0 - 5.0 * std::sqrt(3.0)) / 36, (3 - std::sqrt(3.0)) / 36, (3 + std::sqrt(3.0)) / 36,
                          (9.0 + 5.0 * std::sqrt(3.0)) / 36};
    const double Dv[3] = {(2.0 - std::sqrt(3.0)) / 12, 1.0 / 12, (2.0 + std::sqrt(3.0)) / 12};
    for (int p = 0; p < 8; p++) {
      T.w[p] = 1.0;
      for (int a = 0; a < 8; a++) {
        int k = 0;
        for (int c = 0; c < 3; c++) k += (ns[a][c] == ns[p][c]);
        T.N[p][a] = Nv[k];
        for (int c = 0; c < 3; c++) {
          int kk = 0;
          for (int o = 0; o < 3; o++)
            if (o != c) kk += (ns[a][o] == ns[p][o]);
          const double v = Dv[kk];
          T.dNdr[p][a][c] = ns[a][c] > 0 ? v : -v;
        }
      }
    }
  }
  return T;
}

/* RampFunction::setFunction / getValue (RampFunction.C:40-131) */
double ramp_value(int type, double mn, double mx, double x) {
  if (type == DNLG_RAMP_NONE) return 0;
  const double mm = mx - mn;
  double a0 = 0, a1 = 0;
  switch (type) {
    case DNLG_RAMP_LINEAR: a0 = -2.0 / mm; a1 = 2.0 / mm; break;
    case DNLG_RAMP_SINUS: a1 = kPI / 2; a0 = 1.0 / (1.0 - 2.0 / kPI) / mm; break;
    case DNLG_RAMP_CONSTANT: a0 = 1.0 / mm; break;
    case DNLG_RAMP_ERF: a0 = 5; a1 = std::sqrt(kPI) * mm / 10; break;
    case DNLG_RAMP_SOFT: a0 = 1 / mm; break;
    default: return 0;
  }
  if ((x < mn) || (x > mx)) return 0;
  const double pt = (x - mn) / (mx - mn);
  switch (type) {
    case DNLG_RAMP_LINEAR: return a0 * pt + a1;
    case DNLG_RAMP_SINUS: return a0 * (1.0 - std::sin(a1 * pt));
    case DNLG_RAMP_CONSTANT: return a0;
    case DNLG_RAMP_ERF: return std::exp(-sq(a0 * pt)) / a1;
    case DNLG_RAMP_SOFT: return a0 * (kPI * std::sin(kPI * pt) / 2);
  }
  return 0;
}

}  // namespace

/* ===========================================================================
 * Oracle job: SoA state, reference loop order.
 * ========================================================================= */
struct OracleJob {
  Tables tab;
  int dim = 0, nn = 0, nip = 0;
  int64_t N = 0, NE = 0;
  dnlg_material mat{};
  dnlg_solver_params prm{};
  Law law{};
  double alphaM = 0, beta = 0, gammaCH = 0, omegaS = 0;
  std::vector<dnlg_boundary> bnd;
  std::vector<int64_t> cptr;
  std::vector<int32_t> cidx;
  std::vector<int32_t> conn;
  /* nodes */
  std::vector<double> x, disp, mass;
  std::vector<double> fu[2], fv[2], fa[2];
  int f0 = 0; /* index of field0 */
  std::vector<double> fint;
  /* IPs */
  std::vector<double> detJ, detJ0, radius, dN; /* dN: [ip_global][nn][dim] */
  std::vector<double> sig, eps, epsP, dE, Rr;  /* [ip_global][6 or 9] */
  std::vector<double> pres, ep, dep, sy, gam, gamc, T, eint, einel, rho;
  std::vector<double> elemF; /* [e][nn*dim] */
  /* power iteration state kept between time steps (Model.h:46, Model.C:966):
   * the iterate in the reference's DOF order (node * dim + d) and the last
   * maximum */
  std::vector<double> pi_ev;
  double pi_fmax = 0.0;
  int pi_iterations = 0; /* iterations of the last call */
  /* history files (HistoryFile.h): save times, items, rows taken */
  struct History {
    double start = 0, next = 0, save = 0;
    std::vector<dnlg_history_item> items;
    std::vector<double> rows; /* 1 + items.size() per row */
  };
  std::vector<History> hist;
  std::vector<std::vector<std::pair<int64_t, int>>> node_elems; /* Node::elements, (element, local node) */
  double t = 0, dt = 0;
  int64_t inc = 0;
  bool started = false;
  int status = DNLG_OK;
  int threads = 1;
  std::string err;
  std::atomic<bool> table_range{false}; /* a tabular law was evaluated outside its table */
};

namespace {

void set_err(OracleJob *J, int st, const char *msg) {
  if (J->status == DNLG_OK) {
    J->status = st;
    J->err = msg;
  }
}

/* Element3D::computeJacobian (Element3D.C:64-115), ElementPlane::
 * computeJacobian (ElementPlane.C:46-102) with the axisymmetric radius
 * (ElementAxi.C:46-59).  Returns false on a negative determinant. */
bool element_jacobian(OracleJob *J, int64_t e, bool reference) {
  const Tables &T = J->tab;
  const int32_t *cn = &J->conn[e * J->nn];
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    double Jm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < J->nn; a++) {
      const int64_t nd = cn[a];
      for (int i = 0; i < J->dim; i++)
        for (int j = 0; j < J->dim; j++) Jm[i][j] += T.dNdr[p][a][j] * J->x[i * J->N + nd];
    }
    double det, inv[3][3];
    if (J->dim == 3) {
      const double *d = &Jm[0][0];
      det = d[1] * d[5] * d[6] - d[2] * d[4] * d[6] + d[2] * d[3] * d[7] - d[0] * d[5] * d[7] -
            d[1] * d[3] * d[8] + d[0] * d[4] * d[8];
      if (det < 0.0) return false;
      double *o = &inv[0][0];
      o[0] = (d[4] * d[8] - d[5] * d[7]) / det;
      o[1] = (d[2] * d[7] - d[1] * d[8]) / det;
      o[2] = (d[1] * d[5] - d[2] * d[4]) / det;
      o[3] = (d[5] * d[6] - d[3] * d[8]) / det;
      o[4] = (d[0] * d[8] - d[2] * d[6]) / det;
      o[5] = (d[2] * d[3] - d[0] * d[5]) / det;
      o[6] = (d[3] * d[7] - d[4] * d[6]) / det;
      o[7] = (d[1] * d[6] - d[0] * d[7]) / det;
      o[8] = (d[0] * d[4] - d[1] * d[3]) / det;
    } else {
      det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
      if (det < 0.0) return false;
      inv[0][0] = Jm[1][1] / det;
      inv[0][1] = -Jm[0][1] / det;
      inv[1][0] = -Jm[1][0] / det;
      inv[1][1] = Jm[0][0] / det;
    }
    J->detJ[g] = det;
    double *dn = &J->dN[g * J->nn * J->dim];
    for (int a = 0; a < J->nn; a++)
      for (int j = 0; j < J->dim; j++) {
        double s = 0.0;
        for (int k = 0; k < J->dim; k++) s += T.dNdr[p][a][k] * inv[k][j];
        dn[a * J->dim + j] = s;
      }
    if (T.axi) {
      double r = 0.0;
      for (int a = 0; a < J->nn; a++) r += T.N[p][a] * J->x[cn[a]];
      J->radius[g] = r;
    }
    if (reference) {
      J->detJ0[g] = det;
      if (T.axi) J->detJ0[g] = det * J->radius[g];
    }
  }
  return true;
}

bool compute_jacobian(OracleJob *J, bool reference) {
  bool ok = true;
#pragma omp parallel for schedule(static) num_threads(J->threads) reduction(&& : ok)
  for (int64_t e = 0; e < J->NE; e++) ok = element_jacobian(J, e, reference) && ok;
  if (!ok) set_err(J, DNLG_E_NEG_JACOBIAN, "negative detJ");
  return ok;
}

/* Element::computeMassMatrix (Element.C:304-339) assembled as in
 * Model::computeMassMatrix (Model.C:537-590). */
void compute_mass(OracleJob *J) {
  J->mass.assign(J->N, 0.0);
  std::vector<double> me(J->nn);
  for (int64_t e = 0; e < J->NE; e++) {
    for (int a = 0; a < J->nn; a++) me[a] = 0.0;
    for (int p = 0; p < J->nip; p++) {
      const int64_t g = e * J->nip + p;
      double wdj = J->tab.w[p] * J->detJ[g];
      if (J->tab.axi) wdj *= k2PI * J->radius[g];
      for (int a = 0; a < J->nn; a++) me[a] += J->mat.density * J->tab.N[p][a] * wdj;
    }
    for (int a = 0; a < J->nn; a++) J->mass[J->conn[e * J->nn + a]] += me[a];
  }
}

/* Characteristic lengths: ElQua4NAx::getCharacteristicLength
 * (ElQua4NAx.C:312-321); ElHex8N3D::getCharacteristicLength / getVolume
 * (ElHex8N3D.C:614-684). */
double char_length(const OracleJob *J, int64_t e) {
  if (J->tab.type == DNLG_EL_TET4N_3D) return 0.0; /* ElTet4N3D.h:45 */
  const int32_t *cn = &J->conn[e * J->nn];
  auto X = [&](int a, int c) { return J->x[c * J->N + cn[a]]; };
  if (J->dim == 2) {
    const double x31 = X(2, 0) - X(0, 0);
    const double x24 = X(1, 0) - X(3, 0);
    const double y42 = X(3, 1) - X(1, 1);
    const double y31 = X(2, 1) - X(0, 1);
    return ((x31 * y42 + x24 * y31) / std::sqrt(2.0 * (sq(x24) + sq(y42) + sq(x31) + sq(y31))));
  }
  auto aert = [&](int p1, int p2, int p3, int p4) {
    const double x1 = X(p1, 0), x2 = X(p2, 0), x3 = X(p3, 0), x4 = X(p4, 0);
    const double y1 = X(p1, 1), y2 = X(p2, 1), y3 = X(p3, 1), y4 = X(p4, 1);
    const double z1 = X(p1, 2), z2 = X(p2, 2), z3 = X(p3, 2), z4 = X(p4, 2);
    return std::sqrt(4 * (sq(x1 - x3) + sq(y1 - y3) + sq(z1 - z3)) * (sq(x2 - x4) + sq(y2 - y4) + sq(z2 - z4)) -
                     4 * sq(-(x2 * x3) + x1 * (x2 - x4) + x3 * x4 + y1 * y2 - y2 * y3 - y1 * y4 + y3 * y4 + z1 * z2 -
                            z2 * z3 - z1 * z4 + z3 * z4)) /
           4;
  };
  auto mx3 = [](double i, double j, double k) { return (i > j) ? ((i > k) ? i : k) : ((j > k) ? j : k); };
  double A1 = aert(0, 1, 2, 3), A2 = aert(0, 1, 5, 4), A3 = aert(1, 2, 6, 5);
  double A4 = aert(3, 2, 6, 7), A5 = aert(0, 3, 7, 4), A6 = aert(4, 5, 6, 7);
  A1 = mx3(A1, A2, A3);
  A2 = mx3(A4, A5, A6);
  A1 = (A1 > A2) ? A1 : A2;
  /* volume from three pyramids */
  auto sub = [&](int a, int b, double o[3]) {
    for (int c = 0; c < 3; c++) o[c] = X(a, c) - X(b, c);
  };
  auto cross = [](const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto nrm = [](const double a[3]) { return std::sqrt(sq(a[0]) + sq(a[1]) + sq(a[2])); };
  auto pyramid = [&](int o, int a, int b, int c, int apex) {
    double va[3], vb[3], vc[3], n1[3], n2[3], nn[3], h[3];
    sub(a, o, va);
    sub(b, o, vb);
    sub(c, o, vc);
    cross(va, vb, n1);
    cross(vc, vb, n2);
    const double base = (nrm(n1) + nrm(n2)) / 2.0;
    cross(va, vc, nn);
    const double ln = nrm(nn);
    for (int k = 0; k < 3; k++) nn[k] /= ln;
    sub(apex, o, h);
    double ht = nn[0] * h[0] + nn[1] * h[1] + nn[2] * h[2];
    ht = (ht < 0) ? -ht : ht;
    return base * ht / 3.0;
  };
  double vol = 0.0;
  vol += pyramid(0, 1, 2, 3, 7);
  vol += pyramid(0, 1, 5, 4, 7);
  vol += pyramid(1, 2, 6, 5, 7);
  return vol / A1;
}

/* Element::getElongationWaveSpeed (Element.C:343-359) */
double wave_speed(const OracleJob *J, int64_t e) {
  double rho = 0.0;
  for (int p = 0; p < J->nip; p++) rho += J->rho[e * J->nip + p];
  rho /= J->nip;
  const double nu = J->mat.poisson;
  return std::sqrt((J->mat.young * (1.0 - nu)) / (rho * (1.0 + nu) * (1.0 - 2.0 * nu)));
}

/* Model::computeCourantTimeStep (Model.C:636-670) */
double courant(OracleJob *J) {
  double crit = char_length(J, 0) / wave_speed(J, 0);
  for (int64_t e = 1; e < J->NE; e++) {
    const double ts = char_length(J, e) / wave_speed(J, e);
    if (ts < crit) crit = ts;
  }
  return crit;
}

/* Material::getHookeStiffnessMatrix (Material.C:98-152), threedimensional
 * (6 x 6, rows xx yy zz xy yz xz) and axisymetric (4 x 4, rows rr zz rz tt)
 * cases; the shear diagonal is (1 - 2 nu), as in the reference.  The
 * planeStrain case is the leading 3 x 3 block of the axisymetric one. */
void hooke(const OracleJob *J, double C[6][6]) {
  const double nu = J->mat.poisson, E = J->mat.young;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) C[i][j] = 0.0;
  if (J->dim == 3) {
    C[0][0] = C[1][1] = C[2][2] = 1 - nu;
    C[3][3] = C[4][4] = C[5][5] = 1 - 2 * nu;
    C[0][1] = C[1][0] = C[0][2] = C[2][0] = C[1][2] = C[2][1] = nu;
  } else {
    C[0][0] = C[1][1] = C[3][3] = 1 - nu;
    C[2][2] = 1 - 2 * nu;
    C[1][0] = C[0][1] = C[3][0] = C[0][3] = C[3][1] = C[1][3] = nu;
  }
  const double f = E / ((1 + nu) * (1 - 2 * nu));
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) C[i][j] *= f;
}

/* One integration point's term of Element3D::computeElasticStiffnessMatrix
 * (Element3D.C:303-399) and ElementPlane::computeElasticStiffnessMatrix
 * (ElementPlane.C:182-292), axisymmetric and plane-strain branches: C B,
 * then K += B^T (C B) WxdJ in the reference's term order.  dn: the shape
 * derivatives [node][dir] at the point, WxdJ = weight detJ, Np: the shape
 * functions at the point, R: its radius (axisymmetric). */
void add_point_stiffness(const OracleJob *J, const double C[6][6], const double *dn, double WxdJ, const double *Np,
                         double R, double *K) {
  const int nn = J->nn, dim = J->dim, nd = nn * dim;
  std::vector<double> CB((size_t)6 * nd);
  auto cb = [&](int r, int c) -> double & { return CB[(size_t)r * nd + c]; };
  auto kk = [&](int r, int c) -> double & { return K[(size_t)r * nd + c]; };
  if (dim == 3) {
    for (int i = 0; i < nn; i++) {
      const int I = 3 * i;
      const double dNx = dn[i * 3], dNy = dn[i * 3 + 1], dNz = dn[i * 3 + 2];
      for (int r = 0; r < 6; r++) {
        cb(r, I) = (C[r][0] * dNx + C[r][3] * dNy + C[r][5] * dNz);
        cb(r, I + 1) = (C[r][1] * dNy + C[r][3] * dNx + C[r][4] * dNz);
        cb(r, I + 2) = (C[r][2] * dNz + C[r][4] * dNy + C[r][5] * dNx);
      }
    }
    for (int i = 0; i < nn; i++) {
      const double dNx = dn[i * 3], dNy = dn[i * 3 + 1], dNz = dn[i * 3 + 2];
      for (int j = 0; j < nn; j++) {
        const int I = 3 * i, Jc = 3 * j;
        for (int b = 0; b < 3; b++) {
          kk(I, Jc + b) += (dNx * cb(0, Jc + b) + dNy * cb(3, Jc + b) + dNz * cb(5, Jc + b)) * WxdJ;
          kk(I + 1, Jc + b) += (dNy * cb(1, Jc + b) + dNx * cb(3, Jc + b) + dNz * cb(4, Jc + b)) * WxdJ;
          kk(I + 2, Jc + b) += (dNz * cb(2, Jc + b) + dNy * cb(4, Jc + b) + dNx * cb(5, Jc + b)) * WxdJ;
        }
      }
    }
  } else {
    const bool axi = J->tab.axi;
    if (axi) WxdJ *= k2PI * R;
    for (int i = 0; i < nn; i++) {
      const int I = 2 * i;
      const double dNx = dn[i * 2], dNy = dn[i * 2 + 1];
      const double NoverR = Np[i] / R;
      for (int r = 0; r < 3; r++) {
        cb(r, I) = (C[r][0] * dNx + C[r][2] * dNy);
        cb(r, I + 1) = (C[r][1] * dNy + C[r][2] * dNx);
      }
      if (axi) {
        cb(3, I) = (C[3][0] * dNx + C[3][2] * dNy);
        cb(3, I + 1) = (C[3][1] * dNy + C[3][2] * dNx);
        for (int r = 0; r < 4; r++) cb(r, I) += C[r][3] * NoverR;
      }
    }
    for (int i = 0; i < nn; i++) {
      const double dNx = dn[i * 2], dNy = dn[i * 2 + 1];
      const double NoverR = Np[i] / R;
      for (int j = 0; j < nn; j++) {
        const int I = 2 * i, Jc = 2 * j;
        for (int b = 0; b < 2; b++) {
          kk(I, Jc + b) += (dNx * cb(0, Jc + b) + dNy * cb(2, Jc + b)) * WxdJ;
          kk(I + 1, Jc + b) += (dNy * cb(1, Jc + b) + dNx * cb(2, Jc + b)) * WxdJ;
        }
        if (axi) {
          kk(I, Jc) += cb(3, Jc) * NoverR * WxdJ;
          kk(I, Jc + 1) += cb(3, Jc + 1) * NoverR * WxdJ;
        }
      }
    }
  }
}

/* computeElasticStiffnessMatrix(false): K_e = sum_ip B^T C B w detJ
 * (x 2 pi r) on the Jacobian of the last compute_jacobian.  K is
 * (nn dim) x (nn dim), row-major. */
void element_stiffness(const OracleJob *J, int64_t e, double *K) {
  const int nn = J->nn, dim = J->dim, nd = nn * dim;
  double C[6][6];
  hooke(J, C);
  for (int k = 0; k < nd * nd; k++) K[k] = 0.0;
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    const bool axi = dim == 2 && J->tab.axi;
    add_point_stiffness(J, C, &J->dN[g * nn * dim], J->tab.w[p] * J->detJ[g], J->tab.N[p], axi ? J->radius[g] : 1.0,
                        K);
  }
}

/* computeElasticStiffnessMatrix(true), PowerIterationUnder: one point at the
 * element centre (the under-integration point of ElHex8N3D.C:233-252,
 * ElQua4NAx.C:112-124, ElQua4N2D.C:115-127: weight 8 / 4, shape functions
 * 1/8 / 1/4, derivatives +-1/8 / +-1/4 by the nodes' local signs), its
 * Jacobian, inverse and radius formed on the current coordinates
 * (Element3D::computeUnderJacobian, Element3D.C:118-167;
 * ElementPlane::computeUnderJacobian, ElementPlane.C:104-160).  false for a
 * negative determinant (the reference exits, Model.C:777-798). */
bool element_stiffness_under(const OracleJob *J, int64_t e, double *K) {
  const int nn = J->nn, dim = J->dim, nd = nn * dim;
  /* local coordinate signs of the reference's node numbering */
  static const double sgn8[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  static const double sgn4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  double dNdr[8][3], Np[8];
  for (int a = 0; a < nn; a++) {
    Np[a] = 1.0 / nn;
    for (int k = 0; k < dim; k++) dNdr[a][k] = (nn == 8 ? sgn8[a][k] : sgn4[a][k]) / nn;
  }
  const int32_t *cn = &J->conn[e * nn];
  double Jm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < nn; a++)
    for (int i = 0; i < dim; i++)
      for (int j = 0; j < dim; j++) Jm[i][j] += dNdr[a][j] * J->x[i * J->N + cn[a]];
  double det, inv[3][3];
  if (dim == 3) {
    const double *d = &Jm[0][0];
    det = d[1] * d[5] * d[6] - d[2] * d[4] * d[6] + d[2] * d[3] * d[7] - d[0] * d[5] * d[7] - d[1] * d[3] * d[8] +
          d[0] * d[4] * d[8];
    if (det < 0.0) return false;
    double *o = &inv[0][0];
    o[0] = (d[4] * d[8] - d[5] * d[7]) / det;
    o[1] = (d[2] * d[7] - d[1] * d[8]) / det;
    o[2] = (d[1] * d[5] - d[2] * d[4]) / det;
    o[3] = (d[5] * d[6] - d[3] * d[8]) / det;
    o[4] = (d[0] * d[8] - d[2] * d[6]) / det;
    o[5] = (d[2] * d[3] - d[0] * d[5]) / det;
    o[6] = (d[3] * d[7] - d[4] * d[6]) / det;
    o[7] = (d[1] * d[6] - d[0] * d[7]) / det;
    o[8] = (d[0] * d[4] - d[1] * d[3]) / det;
  } else {
    det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
    if (det < 0.0) return false;
    inv[0][0] = Jm[1][1] / det;
    inv[0][1] = -Jm[0][1] / det;
    inv[1][0] = -Jm[1][0] / det;
    inv[1][1] = Jm[0][0] / det;
  }
  double dn[8 * 3];
  for (int a = 0; a < nn; a++)
    for (int j = 0; j < dim; j++) {
      double s = 0.0;
      for (int k = 0; k < dim; k++) s += dNdr[a][k] * inv[k][j];
      dn[a * dim + j] = s;
    }
  double R = 1.0;
  if (dim == 2 && J->tab.axi) {
    R = 0.0;
    for (int a = 0; a < nn; a++) R += Np[a] * J->x[cn[a]];
  }
  double C[6][6];
  hooke(J, C);
  for (int k = 0; k < nd * nd; k++) K[k] = 0.0;
  add_point_stiffness(J, C, dn, (double)nn * det, Np, R, K);
  return true;
}

/* Model::computePowerIterationTimeStep (Model.C:942-1009): element
 * stiffness matrices (full integration, or for PowerIterationUnder the
 * one-point under-integration), then iterations of
 * v <- M^-1 (v + sum_e K_e v_e) (Vector::gatherFrom accumulates into the
 * iterate, Vector.C:893-945; MatrixDiag::divideBy, MatrixDiag.C:655-670),
 * fmax = max |v|, v <- v / fmax, until |fmax - fmax_prev| / fmax < 1e-4
 * (Model.h:46-48).  The iterate and fmax_prev persist between calls; the
 * first call starts from v_0 = 1, v_i = v_{i-1} - 2 / (n - 1).  Returns the
 * maximum frequency sqrt(fmax), a negative value after 1000 iterations, or
 * 0 for a negative under-integration Jacobian. */
double power_iteration(OracleJob *J) {
  const int nd = J->nn * J->dim;
  const int64_t ndof = (int64_t)J->dim * J->N;
  std::vector<double> Ke((size_t)J->NE * nd * nd);
  if (J->prm.dt_method == DNLG_DT_POWER_ITERATION_UNDER && J->nip > 1) {
    bool ok = true;
#pragma omp parallel for schedule(static) num_threads(J->threads) reduction(&& : ok)
    for (int64_t e = 0; e < J->NE; e++) ok = element_stiffness_under(J, e, &Ke[(size_t)e * nd * nd]) && ok;
    if (!ok) return 0.0;
  } else {
#pragma omp parallel for schedule(static) num_threads(J->threads)
    for (int64_t e = 0; e < J->NE; e++) element_stiffness(J, e, &Ke[(size_t)e * nd * nd]);
  }
  if (J->pi_fmax == 0.0 || (int64_t)J->pi_ev.size() != ndof) {
    J->pi_ev.assign(ndof, 0.0);
    J->pi_ev[0] = 1.0;
    for (int64_t i = 1; i < ndof; i++) J->pi_ev[i] = J->pi_ev[i - 1] - 2. / (ndof - 1);
  }
  std::vector<double> &ev = J->pi_ev;
  std::vector<double> ev0, loc(nd), y(nd);
  double fmax = 0.0;
  int iteration = 0;
  while (true) {
    iteration++;
    ev0 = ev;
    for (int64_t e = 0; e < J->NE; e++) {
      const int32_t *cn = &J->conn[e * J->nn];
      for (int a = 0; a < J->nn; a++)
        for (int d = 0; d < J->dim; d++) loc[a * J->dim + d] = ev0[(int64_t)cn[a] * J->dim + d];
      const double *K = &Ke[(size_t)e * nd * nd];
      for (int r = 0; r < nd; r++) {
        double sum = 0.0;
        for (int c = 0; c < nd; c++) sum += K[(size_t)r * nd + c] * loc[c];
        y[r] = sum;
      }
      for (int a = 0; a < J->nn; a++)
        for (int d = 0; d < J->dim; d++) ev[(int64_t)cn[a] * J->dim + d] += y[a * J->dim + d];
    }
    for (int64_t i = 0; i < ndof; i++) ev[i] /= J->mass[i / J->dim];
    fmax = std::fabs(ev[0]);
    for (int64_t i = 1; i < ndof; i++)
      if (std::fabs(ev[i]) > fmax) fmax = std::fabs(ev[i]);
    for (int64_t i = 0; i < ndof; i++) ev[i] /= fmax;
    const double convergence = std::fabs(fmax - J->pi_fmax) / fmax;
    const bool ok = convergence < 1e-4;
    J->pi_fmax = fmax;
    if (iteration > 1000) { /* checked before the exit, as Model.C:1002 */
      J->pi_iterations = iteration;
      return -1.0;
    }
    if (ok) break;
  }
  J->pi_iterations = iteration;
  return std::sqrt(fmax);
}

/* Solver::computeTimeStep (Solver.C:196-236): Courant (omega_max =
 * 2 / dt_crit) or power-iteration branch */
bool compute_time_step(OracleJob *J, bool force) {
  if (force || (J->inc % J->prm.dt_freq == 0)) {
    double fmax;
    if (J->prm.dt_method == DNLG_DT_POWER_ITERATION || J->prm.dt_method == DNLG_DT_POWER_ITERATION_UNDER) {
      fmax = power_iteration(J);
      if (fmax == 0.0) {
        set_err(J, DNLG_E_NEG_JACOBIAN, "negative detJ at an under-integration point");
        return false;
      }
      if (!(fmax > 0.0)) {
        set_err(J, DNLG_E_PI_DIVERGED, "power iteration not converged in 1000 iterations");
        return false;
      }
    } else {
      fmax = 2.0 / courant(J);
    }
    J->dt = J->prm.safety * J->omegaS / fmax;
  }
  return true;
}

/* Element::computeStrains (Element.C:1103-1122) with the deformation
 * gradient of Element3D.C:172-195 / ElementAxi.C:78-100, built from the
 * field0 displacement increment. */
void element_strains(OracleJob *J, int64_t e) {
  const int32_t *cn = &J->conn[e * J->nn];
  const std::vector<double> &u = J->fu[J->f0];
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    const double *dn = &J->dN[g * J->nn * J->dim];
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double vr = 0.0;
    for (int a = 0; a < J->nn; a++) {
      const int64_t nd = cn[a];
      for (int i = 0; i < J->dim; i++)
        for (int j = 0; j < J->dim; j++) F[3 * i + j] += dn[a * J->dim + j] * u[i * J->N + nd];
      if (J->tab.axi) vr += J->tab.N[p][a] * u[nd];
    }
    if (J->tab.axi) F[8] += vr / J->radius[g];
    polar_lnU(F, &J->dE[g * 6], &J->Rr[g * 9]);
    if (J->prm.track_tensors)
      for (int c = 0; c < 6; c++) J->eps[g * 6 + c] += J->dE[g * 6 + c];
  }
}

/* Element::computePressure (Element.C:1125-1143) */
void element_pressure(OracleJob *J, int64_t e) {
  const double K = J->mat.young / (3.0 * (1.0 - 2.0 * J->mat.poisson));
  double inc = 0.0;
  for (int p = 0; p < J->nip; p++) {
    const double *d = &J->dE[(e * J->nip + p) * 6];
    inc += (d[0] + d[3] + d[5]);
  }
  inc /= J->nip;
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    J->pres[g] = third_trace(&J->sig[g * 6]) + K * inc;
  }
}

/* Element::computeStress (Element.C:407-622): J2 radial return with a
 * safeguarded Newton-Raphson on gamma, adiabatic heating, energies. */
bool element_stress(OracleJob *J, int64_t e, double timeStep) {
  const double gammaInitial = 1e-8, tolNR = 1e-8;
  const int itMax = J->prm.nr_max_iter > 0 ? J->prm.nr_max_iter : 250; /* Element.C:560-581 */
  const double TwoG = J->mat.young / (1.0 + J->mat.poisson);
  const double TwoG32 = kSqrt32 * TwoG;
  const double heatFrac = J->mat.taylor_quinney / (J->mat.density * J->mat.heat_capacity);
  const Law &L = J->law;
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    double old[6], s[6], dd[6];
    for (int c = 0; c < 6; c++) old[c] = J->sig[g * 6 + c];
    deviator(old, s);
    const double Snorm0 = norm_sym(s);
    deviator(&J->dE[g * 6], dd);
    for (int c = 0; c < 6; c++) s[c] += TwoG * dd[c];
    const double Snorm = norm_sym(s);
    const double Strial = kSqrt32 * Snorm;
    const double T0 = J->T[g];
    double T = T0;
    double epsp = J->ep[g];
    double depsp = J->dep[g];
    double gamma = 0.0;
    double yield = J->sy[g];
    if (yield == 0.0) {
      yield = law_yield(L, gammaInitial, gammaInitial / timeStep, T);
      J->sy[g] = yield;
    }
    if (std::isnan(yield)) {
      J->table_range = true;
      return false;
    }
    int iterate = 0;
    if (Strial > yield && J->prm.stress_method == DNLG_STRESS_DIRECT) {
      /* Element::computeStressDirect (Element.C:625-768): one linearised
       * return from the stored yield stress and the slope at the stored
       * state (plastic strain, or gammaInitial before the first yield) */
      const double hard = law_dyield(L, epsp > 0.0 ? epsp : gammaInitial, depsp, T, timeStep);
      gamma = (Snorm - kSqrt23 * yield) / (TwoG * (1.0 + hard / (1.5 * TwoG)));
      depsp = kSqrt23 * gamma / timeStep;
      epsp = J->ep[g] + kSqrt23 * gamma;
      T = T0 + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
      yield = law_yield(L, epsp, depsp, T);
      if (std::isnan(yield) || std::isnan(hard)) {
        J->table_range = true;
        return false;
      }
      if (J->prm.track_tensors)
        for (int c = 0; c < 6; c++) J->epsP[g * 6 + c] += gamma * s[c] / Snorm;
      const double scale = 1.0 - TwoG * gamma / Snorm;
      for (int c = 0; c < 6; c++) s[c] *= scale;
      J->ep[g] += kSqrt23 * gamma;
      J->dep[g] = kSqrt23 * gamma / timeStep;
      J->gam[g] = gamma;
      J->gamc[g] += gamma;
      J->sy[g] = yield;
    } else if (Strial > yield) {
      double gmin = 0.0;
      double gmax = (Strial - yield) / TwoG32;
      gamma = J->gam[g];
      if (epsp == 0.0) gamma = kSqrt32 * gammaInitial;
      depsp = kSqrt23 * gamma / timeStep;
      epsp = J->ep[g] + kSqrt23 * gamma;
      T = T0 + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
      bool run = true;
      while (run) {
        yield = law_yield(L, epsp, depsp, T);
        if (std::isnan(yield)) {
          J->table_range = true;
          return false;
        }
        const double fun = Strial - gamma * TwoG32 - yield;
        if (fun < 0.0)
          gmax = gamma;
        else
          gmin = gamma;
        const double hard = law_dyield(L, epsp, depsp, T, timeStep);
        const double dfun = TwoG32 + kSqrt23 * hard;
        double dgamma = fun / dfun;
        gamma += dgamma;
        if ((gmax - gamma) * (gamma - gmin) < 0.0) {
          dgamma = 0.5 * (gmax - gmin);
          gamma = gmin + dgamma;
        }
        if (std::fabs(dgamma) < tolNR) {
          run = false;
        } else {
          depsp = kSqrt23 * gamma / timeStep;
          epsp = J->ep[g] + kSqrt23 * gamma;
          T = T0 + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
          iterate += 1;
          if (iterate > itMax) return false;
        }
      }
      if (J->prm.track_tensors)
        for (int c = 0; c < 6; c++) J->epsP[g * 6 + c] += gamma * s[c] / Snorm;
      const double scale = 1.0 - TwoG * gamma / Snorm;
      for (int c = 0; c < 6; c++) s[c] *= scale;
      J->ep[g] += kSqrt23 * gamma;
      J->dep[g] = kSqrt23 * gamma / timeStep;
      J->gam[g] = gamma;
      J->gamc[g] += gamma;
      J->sy[g] = yield;
    }
    double nw[6];
    for (int c = 0; c < 6; c++) nw[c] = s[c];
    nw[0] += J->pres[g];
    nw[3] += J->pres[g];
    nw[5] += J->pres[g];
    for (int c = 0; c < 6; c++) J->sig[g * 6 + c] = nw[c];
    double sum[6];
    for (int c = 0; c < 6; c++) sum[c] = old[c] + nw[c];
    const double power = 0.5 * ddot_sym(&J->dE[g * 6], sum);
    J->eint[g] += power / J->mat.density;
    if (gamma != 0.0) {
      const double w = 0.5 * gamma * (norm_sym(s) + Snorm0);
      J->einel[g] += w / J->mat.density;
      J->T[g] += heatFrac * w;
    }
  }
  return true;
}

/* Element::computeFinalRotation (Element.C:1146-1161) */
void element_rotation(OracleJob *J, int64_t e) {
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    rotate_sym(&J->Rr[g * 9], &J->sig[g * 6]);
    if (J->prm.track_tensors) {
      rotate_sym(&J->Rr[g * 9], &J->eps[g * 6]);
      rotate_sym(&J->Rr[g * 9], &J->epsP[g * 6]);
    }
  }
}

/* Element::computeInternalForces (Element.C:362-404) */
void element_forces(OracleJob *J, int64_t e) {
  double *f = &J->elemF[e * J->nn * J->dim];
  for (int k = 0; k < J->nn * J->dim; k++) f[k] = 0.0;
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    const double *dn = &J->dN[g * J->nn * J->dim];
    const double *S = &J->sig[g * 6];
    double wdj = J->tab.w[p] * J->detJ[g];
    if (J->tab.axi) wdj *= k2PI * J->radius[g];
    for (int a = 0; a < J->nn; a++) {
      for (int i = 0; i < J->dim; i++)
        for (int j = 0; j < J->dim; j++) f[a * J->dim + i] -= dn[a * J->dim + j] * S[sym_idx(j, i)] * wdj;
      if (J->tab.axi) f[a * J->dim] -= J->tab.N[p][a] * S[5] / J->radius[g] * wdj;
    }
  }
}

/* Element::computeDensity (Element.C:1243-1256) */
void element_density(OracleJob *J, int64_t e) {
  for (int p = 0; p < J->nip; p++) {
    const int64_t g = e * J->nip + p;
    J->rho[g] = J->mat.density * J->detJ0[g] / J->detJ[g];
    if (J->tab.axi) J->rho[g] /= J->radius[g];
  }
}

/* Model::computeInternalForces (Model.C:864-899): element-order assembly */
void assemble_forces(OracleJob *J) {
  J->fint.assign(J->dim * J->N, 0.0);
  for (int64_t e = 0; e < J->NE; e++)
    for (int a = 0; a < J->nn; a++) {
      const int64_t gl = (int64_t)J->conn[e * J->nn + a] * J->dim;
      for (int d = 0; d < J->dim; d++) J->fint[gl + d] += J->elemF[(e * J->nn + a) * J->dim + d];
    }
}

/* the five element phases strains .. forces of one cycle */
bool element_phases(OracleJob *J, double timeStep) {
  bool ok = true;
#pragma omp parallel for schedule(static) num_threads(J->threads) reduction(&& : ok)
  for (int64_t e = 0; e < J->NE; e++) {
    element_strains(J, e);
    element_pressure(J, e);
    ok = element_stress(J, e, timeStep) && ok;
    element_rotation(J, e);
    element_forces(J, e);
  }
  if (!ok) {
    if (J->table_range)
      set_err(J, DNLG_E_TABLE_RANGE, "plastic strain outside the tabular law");
    else
      set_err(J, DNLG_E_NR_DIVERGED, "no convergence in Newton-Raphson");
    return false;
  }
  assemble_forces(J);
  return true;
}

/* BoundaryRestrain / BoundarySpeed ::applyConstantOnNewFields
 * (BoundaryRestrain.C:98-150, BoundarySpeed.C:139-226), node's constant
 * list in attach order (BoundaryCondition.C:174-190). */
void apply_constant_bc(OracleJob *J, int64_t n, double time, double timeStep) {
  if (J->cptr.empty()) return;
  const int a = J->f0, b = 1 - J->f0;
  for (int64_t k = J->cptr[n]; k < J->cptr[n + 1]; k++) {
    const dnlg_boundary &B = J->bnd[J->cidx[k]];
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      if (B.kind == DNLG_BC_RESTRAIN) {
        if (B.value[d] > 0) {
          J->fu[a][q] = 0.0;
          J->fu[b][q] = 0.0;
          J->fv[a][q] = 0.0;
          J->fv[b][q] = 0.0;
          J->fa[a][q] = 0.0;
          J->fa[b][q] = 0.0;
        }
      } else if (B.value[d] != 0.0) {
        if (B.ramp == DNLG_RAMP_NONE) {
          J->fa[a][q] = 0.0;
          J->fa[b][q] = 0.0;
          J->fv[a][q] = B.value[d];
          J->fv[b][q] = B.value[d];
          J->fu[a][q] = J->fv[a][q] * timeStep;
          J->fu[b][q] = J->fv[b][q] * timeStep;
        } else {
          const double f1 = ramp_value(B.ramp, B.ramp_min, B.ramp_max, time + timeStep);
          const double f0 = ramp_value(B.ramp, B.ramp_min, B.ramp_max, time);
          J->fa[a][q] = (f1 - f0) * B.value[d] / timeStep;
          J->fa[b][q] = J->fa[a][q];
          J->fv[a][q] = f0 * B.value[d];
          J->fv[b][q] = f1 * B.value[d];
          J->fu[a][q] = J->fv[a][q] * timeStep;
          J->fu[b][q] = J->fv[b][q] * timeStep;
        }
      }
    }
  }
}

/* Explicit::computePredictions (Explicit.C:438-482) */
void predictor(OracleJob *J) {
  const int a = J->f0, b = 1 - J->f0;
  const double h = J->dt;
  for (int64_t n = 0; n < J->N; n++) {
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      J->fu[b][q] = h * (J->fv[a][q] + (0.5 - J->beta) * h * J->fa[a][q]);
      J->fv[b][q] = J->fv[a][q] + (1.0 - J->gammaCH) * h * J->fa[a][q];
      J->fa[b][q] = 0.0;
    }
    apply_constant_bc(J, n, J->t, h);
  }
}

/* Explicit::explicitSolve (Explicit.C:489-536) with MatrixDiag::solve
 * (MatrixDiag.C:751-756) */
void corrector(OracleJob *J) {
  const int a = J->f0, b = 1 - J->f0;
  const double h = J->dt;
  for (int64_t n = 0; n < J->N; n++) {
    for (int d = 0; d < J->dim; d++) J->fa[b][d * J->N + n] = J->fint[n * J->dim + d] / J->mass[n];
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      J->fa[b][q] -= J->alphaM * J->fa[a][q];
      J->fa[b][q] /= (1.0 - J->alphaM);
    }
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      J->fv[b][q] += J->gammaCH * h * J->fa[b][q];
    }
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      J->fu[b][q] += J->beta * sq(h) * J->fa[b][q];
    }
    apply_constant_bc(J, n, J->t, h);
    for (int d = 0; d < 3; d++) {
      const int64_t q = d * J->N + n;
      J->disp[q] += J->fu[b][q];
      J->x[q] += J->fu[b][q];
    }
  }
}

void densities(OracleJob *J) {
#pragma omp parallel for schedule(static) num_threads(J->threads)
  for (int64_t e = 0; e < J->NE; e++) element_density(J, e);
}

/* the coordinates and the speeds of field `f` are finite */
bool nodes_finite(const OracleJob *J, int f) {
  for (int64_t q = 0; q < 3 * J->N; q++)
    if (!std::isfinite(J->x[q]) || !std::isfinite(J->fv[f][q])) return false;
  return true;
}

/* The integrationPointsToNode vectors of ElQua4NAx.C:30-60 and
 * ElHex8N3D.C:35-90, [node][ip]: the entry depends only on k, the number of
 * directions in which IP and node lie on the same side of the element
 * centre (hex: k = 3 -> (5 + 3 sqrt3)/4, 2 -> -(1 + sqrt3)/4,
 * 1 -> (sqrt3 - 1)/4, 0 -> (5 - 3 sqrt3)/4; quad: 2 -> 1 + sqrt3/2,
 * 1 -> -1/2, 0 -> 1 - sqrt3/2). */
void extrapolation(int type, double E[8][8]) {
  const double r3 = std::sqrt(3.0);
  if (type == DNLG_EL_TET4N_3D) {
    /* ElTet4N3D.C:33-60: each node's vector has four entries (those of a
     * four-point rule) while the element has one integration point, so the
     * first entry is the one used: (1 + 3 sqrt5)/4 at node 1, (1 - sqrt5)/4
     * at the others */
    const double r5 = std::sqrt(5.0);
    for (int a = 0; a < 4; a++) E[a][0] = a == 0 ? (1.0 + 3.0 * r5) / 4 : (1.0 - r5) / 4;
    return;
  }
  if (type == DNLG_EL_HEX8N_3D) {
    const int ns[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                          {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    const double v[4] = {(5.0 - 3.0 * r3) / 4, (-1.0 + r3) / 4, (-1.0 - r3) / 4, (5.0 + 3.0 * r3) / 4};
    for (int a = 0; a < 8; a++)
      for (int p = 0; p < 8; p++) {
        int k = 0;
        for (int c = 0; c < 3; c++) k += ns[a][c] == ns[p][c];
        E[a][p] = v[k];
      }
  } else {
    const int ns[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    const double v[3] = {1.0 - r3 / 2, -0.5, 1.0 + r3 / 2};
    for (int a = 0; a < 4; a++)
      for (int p = 0; p < 4; p++) {
        int k = 0;
        for (int c = 0; c < 2; c++) k += ns[a][c] == ns[p][c];
        E[a][p] = v[k];
      }
  }
}

/* Element::getIntPointValueExtract (Element.C:1195-1221) at IP g: the
 * field's value, a tensor's component or norm (SymTensor2::norm), 0 for a
 * field the function does not know.  Tensor blocks are norm, XX, XY, XZ,
 * YX, YY, YZ, ZX, ZY, ZZ. */
double ip_field(const OracleJob *J, int f, int64_t g) {
  switch (f) {
    case DNLG_FIELD_PLASTIC_STRAIN: return J->ep[g];
    case DNLG_FIELD_PLASTIC_STRAIN_RATE: return J->dep[g];
    case DNLG_FIELD_GAMMA: return J->gam[g];
    case DNLG_FIELD_GAMMA_CUMULATE: return J->gamc[g];
    case DNLG_FIELD_YIELD_STRESS: return J->sy[g];
    case DNLG_FIELD_T: return J->T[g];
    case DNLG_FIELD_DENSITY: return J->rho[g];
    case DNLG_FIELD_PRESSURE: return J->pres[g];
    case DNLG_FIELD_INTERNAL_ENERGY: return J->eint[g];
    case DNLG_FIELD_VON_MISES: {
      const double *s = &J->sig[g * 6];
      return std::sqrt((sq(s[0] - s[3]) + sq(s[0] - s[5]) + sq(s[3] - s[5]) + 6 * (sq(s[1]) + sq(s[2]) + sq(s[4]))) /
                       2);
    }
  }
  const std::vector<double> *t = nullptr;
  int base = 0;
  if (f >= DNLG_FIELD_STRESS && f < DNLG_FIELD_STRESS + 10) {
    t = &J->sig;
    base = DNLG_FIELD_STRESS;
  } else if (f >= DNLG_FIELD_STRAIN && f < DNLG_FIELD_STRAIN + 10) {
    t = &J->eps;
    base = DNLG_FIELD_STRAIN;
  } else if (f >= DNLG_FIELD_PLASTIC_STRAIN_TENSOR && f < DNLG_FIELD_PLASTIC_STRAIN_TENSOR + 10) {
    t = &J->epsP;
    base = DNLG_FIELD_PLASTIC_STRAIN_TENSOR;
  }
  if (!t || t->empty()) return 0.0;
  const double *s = &(*t)[g * 6];
  const int c = f - base;
  if (c == 0) return norm_sym(s);
  const int rows[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2}, cols[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  return s[sym_idx(rows[c - 1], cols[c - 1])];
}

bool is_ip_field(int f) {
  switch (f) {
    case DNLG_FIELD_PLASTIC_STRAIN: case DNLG_FIELD_PLASTIC_STRAIN_RATE: case DNLG_FIELD_GAMMA:
    case DNLG_FIELD_GAMMA_CUMULATE: case DNLG_FIELD_YIELD_STRESS: case DNLG_FIELD_T: case DNLG_FIELD_DENSITY:
    case DNLG_FIELD_PRESSURE: case DNLG_FIELD_INTERNAL_ENERGY: case DNLG_FIELD_VON_MISES:
      return true;
  }
  return (f >= DNLG_FIELD_STRESS && f < DNLG_FIELD_STRESS + 10) ||
         (f >= DNLG_FIELD_STRAIN && f < DNLG_FIELD_STRAIN + 10) ||
         (f >= DNLG_FIELD_PLASTIC_STRAIN_TENSOR && f < DNLG_FIELD_PLASTIC_STRAIN_TENSOR + 10);
}

/* Node::fieldScalar (Node.C:311-370) */
double node_value(const OracleJob *J, int f, int64_t n) {
  if (f == DNLG_FIELD_MASS) return J->mass[n];
  if (f < DNLG_FIELD_DENSITY) {
    const std::vector<double> *v[5] = {&J->x, &J->disp, &J->fu[J->f0], &J->fv[J->f0], &J->fa[J->f0]};
    const std::vector<double> &a = *v[f / 4];
    if (f % 4 == 0) return std::sqrt(sq(a[n]) + sq(a[J->N + n]) + sq(a[2 * J->N + n]));
    return a[(f % 4 - 1) * J->N + n];
  }
  if (!is_ip_field(f)) return 0.0;
  double E[8][8];
  extrapolation(J->tab.type, E);
  const auto &els = J->node_elems[n];
  const bool tensor = f == DNLG_FIELD_STRESS || f == DNLG_FIELD_STRAIN || f == DNLG_FIELD_PLASTIC_STRAIN_TENSOR;
  if (tensor) {
    /* the norm of the extrapolated tensor */
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const std::vector<double> &t = f == DNLG_FIELD_STRESS ? J->sig : (f == DNLG_FIELD_STRAIN ? J->eps : J->epsP);
    for (const auto &el : els)
      for (int p = 0; p < J->nip; p++)
        for (int c = 0; c < 6; c++) acc[c] += E[el.second][p] * t[(el.first * J->nip + p) * 6 + c];
    for (int c = 0; c < 6; c++) acc[c] = acc[c] / (double)els.size();
    return norm_sym(acc);
  }
  double v = 0.0;
  for (const auto &el : els)
    for (int p = 0; p < J->nip; p++) v += E[el.second][p] * ip_field(J, f, el.first * J->nip + p);
  return v / (double)els.size();
}

/* HistoryFile*Item::getValue (HistoryFile.C:46-99) */
double item_value(const OracleJob *J, const dnlg_history_item &it) {
  if (it.kind == DNLG_H_NODE) return node_value(J, it.field, it.index);
  if (it.kind == DNLG_H_ELEMENT) {
    if (!is_ip_field(it.field)) return 0.0;
    /* Element::getIntPointValue (Element.C:1224-1238) */
    if (it.int_pt > 0) return ip_field(J, it.field, it.index * J->nip + it.int_pt - 1);
    double v = 0.0;
    for (int p = 0; p < J->nip; p++) v += ip_field(J, it.field, it.index * J->nip + p);
    return v / J->nip;
  }
  switch (it.field) {
    case DNLG_FIELD_REAL_TIME_STEP:
    case DNLG_FIELD_TIME_STEP: return J->dt;
    case DNLG_FIELD_KINETIC_ENERGY: {
      /* Model::getTotalKineticEnergy (Model.C:612-625) */
      double ke = 0.0;
      for (int64_t n = 0; n < J->N; n++) {
        double v2 = 0.0;
        for (int d = 0; d < 3; d++) v2 += sq(J->fv[J->f0][d * J->N + n]);
        ke += (J->mass[n] * v2) / 2.0;
      }
      return ke;
    }
  }
  return 0.0;
}

/* Model::writeHistoryFiles -> HistoryFile::save (Model.C:917-924,
 * HistoryFile.C:283-315) */
void write_history(OracleJob *J) {
  for (auto &h : J->hist) {
    if ((J->t > h.start) && (J->t < h.next)) continue;
    h.next += h.save;
    h.rows.push_back(J->t);
    for (const auto &it : h.items) h.rows.push_back(item_value(J, it));
  }
}

}  // namespace

/* ===========================================================================
 * C API of the oracle (mirrors the engine's dnlg_* entry points).
 * ========================================================================= */
extern "C" {

int oracle_job_create(const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bcs *bcs,
                      const dnlg_solver_params *prm, OracleJob **out) {
  if (!mesh || !mat || !prm || !out) return DNLG_E_INVALID;
  if (mesh->elem_type != DNLG_EL_QUA4N_AX && mesh->elem_type != DNLG_EL_HEX8N_3D &&
      mesh->elem_type != DNLG_EL_QUA4N_2D && mesh->elem_type != DNLG_EL_TET4N_3D)
    return DNLG_E_INVALID;
  if (mesh->n_nodes <= 0 || mesh->n_elems <= 0 || !mesh->coords || !mesh->conn) return DNLG_E_INVALID;
  /* ElTet4N3D::getCharacteristicLength returns 0 (ElTet4N3D.h:45): its
   * Courant step is 0 */
  if (mesh->elem_type == DNLG_EL_TET4N_3D && prm->dt_method == DNLG_DT_COURANT) return DNLG_E_INVALID;
  if (prm->dt_method != DNLG_DT_COURANT && prm->dt_method != DNLG_DT_POWER_ITERATION &&
      prm->dt_method != DNLG_DT_POWER_ITERATION_UNDER)
    return DNLG_E_INVALID;
  if (prm->stress_method != DNLG_STRESS_NR && prm->stress_method != DNLG_STRESS_DIRECT) return DNLG_E_INVALID;
  if (prm->dt_freq <= 0) return DNLG_E_INVALID;
  OracleJob *J = new OracleJob();
  J->tab = make_tables(mesh->elem_type);
  J->dim = J->tab.dim;
  J->nn = J->tab.nn;
  J->nip = J->tab.nip;
  J->N = mesh->n_nodes;
  J->NE = mesh->n_elems;
  J->mat = *mat;
  J->prm = *prm;
  J->law = law_of(mat);
  /* Explicit::computeChungHulbertIntegrationParameters (Explicit.C:86-104) */
  const double rb = prm->rho_b;
  J->alphaM = (2.0 * rb - 1.0) / (1.0 + rb);
  J->beta = (5.0 - 3.0 * rb) / ((1.0 + rb) * (1.0 + rb) * (2.0 - rb));
  J->gammaCH = 1.5 - J->alphaM;
  J->omegaS = std::sqrt((12.0 * std::pow(1.0 + rb, 3) * (2.0 - rb)) /
                        (10.0 + 15.0 * rb - rb * rb + std::pow(rb, 3) - std::pow(rb, 4)));
  J->conn.assign(mesh->conn, mesh->conn + J->NE * J->nn);
  for (int64_t k = 0; k < J->NE * J->nn; k++)
    if (J->conn[k] < 0 || J->conn[k] >= J->N) {
      delete J;
      return DNLG_E_INVALID;
    }
  J->x.assign(3 * J->N, 0.0);
  for (int d = 0; d < 3; d++)
    if (d < J->dim)
      for (int64_t n = 0; n < J->N; n++) J->x[d * J->N + n] = mesh->coords[d * J->N + n];
  /* ElementAxi::checkLevel2 (ElementAxi.C:103-126); ElQua4N2D checks nothing */
  if (J->tab.axi)
    for (int64_t n = 0; n < J->N; n++)
      if (J->x[n] < 0.0) {
        if (J->x[n] > -1e-8)
          J->x[n] = 0.0;
        else {
          delete J;
          return DNLG_E_INVALID;
        }
      }
  J->disp.assign(3 * J->N, 0.0);
  for (int k = 0; k < 2; k++) {
    J->fu[k].assign(3 * J->N, 0.0);
    J->fv[k].assign(3 * J->N, 0.0);
    J->fa[k].assign(3 * J->N, 0.0);
  }
  const int64_t G = J->NE * J->nip;
  J->detJ.assign(G, 0.0);
  J->detJ0.assign(G, 0.0);
  J->radius.assign(G, 1.0);
  J->dN.assign(G * J->nn * J->dim, 0.0);
  J->sig.assign(G * 6, 0.0);
  J->dE.assign(G * 6, 0.0);
  J->Rr.assign(G * 9, 0.0);
  if (prm->track_tensors) {
    J->eps.assign(G * 6, 0.0);
    J->epsP.assign(G * 6, 0.0);
  }
  J->pres.assign(G, 0.0);
  J->ep.assign(G, 0.0);
  J->dep.assign(G, 0.0);
  J->sy.assign(G, 0.0);
  J->gam.assign(G, 0.0);
  J->gamc.assign(G, 0.0);
  J->T.assign(G, mat->T0);      /* Element::initializeData (Element.C:258-301) */
  J->rho.assign(G, mat->density);
  J->eint.assign(G, 0.0);
  J->einel.assign(G, 0.0);
  J->elemF.assign(J->NE * J->nn * J->dim, 0.0);
  /* boundary conditions */
  if (bcs && bcs->n_boundaries > 0) {
    J->bnd.assign(bcs->boundaries, bcs->boundaries + bcs->n_boundaries);
    if (bcs->constant_ptr) {
      J->cptr.assign(bcs->constant_ptr, bcs->constant_ptr + J->N + 1);
      J->cidx.assign(bcs->constant_idx, bcs->constant_idx + J->cptr[J->N]);
    }
    /* BoundaryCondition::applyInitial (BoundaryCondition.C:131-150):
     * initial list then constant list; BoundarySpeed::applyInitial
     * (BoundarySpeed.C:63-76) sets the whole speed vector; restrain is a
     * no-op at this stage (BoundaryRestrain.C:60-64). */
    for (int64_t n = 0; n < J->N; n++) {
      auto apply = [&](const int64_t *ptr, const int32_t *idx) {
        if (!ptr) return;
        for (int64_t k = ptr[n]; k < ptr[n + 1]; k++) {
          const dnlg_boundary &B = bcs->boundaries[idx[k]];
          if (B.kind != DNLG_BC_SPEED) continue;
          for (int d = 0; d < 3; d++) {
            J->fv[0][d * J->N + n] = B.value[d];
            J->fv[1][d * J->N + n] = B.value[d];
            J->fa[0][d * J->N + n] = 0.0;
            J->fa[1][d * J->N + n] = 0.0;
          }
        }
      };
      apply(bcs->initial_ptr, bcs->initial_idx);
      apply(bcs->constant_ptr, bcs->constant_idx);
    }
  }
  J->t = prm->t_start;
  *out = J;
  return DNLG_OK;
}

int oracle_job_destroy(OracleJob *J) {
  delete J;
  return DNLG_OK;
}

int oracle_job_set_threads(OracleJob *J, int threads) {
  J->threads = threads < 1 ? 1 : threads;
  return DNLG_OK;
}

/* Explicit::solve (Explicit.C:126-253).  The pre-loop block runs on the
 * first call only. */
int oracle_job_solve(OracleJob *J, double upTo) {
  if (J->status != DNLG_OK) return J->status;
  if (upTo > J->prm.t_end) upTo = J->prm.t_end;
  if (!J->started) {
    J->started = true;
    if (!compute_jacobian(J, true)) return J->status;
    compute_mass(J);
    if (!compute_time_step(J, true)) return J->status;
    if (!element_phases(J, J->dt)) return J->status;
    write_history(J);
  } else if (J->t < upTo) {
    /* A later call continues the loop (the engine's semantics; the
     * reference re-runs its pre-loop block instead).  The last cycle of the
     * previous call skipped the end-of-cycle Jacobian (Explicit.C:243-247),
     * so it is computed here on the current coordinates, as a single call
     * would have done. */
    if (!compute_jacobian(J, false)) return J->status;
  }
  while (J->t < upTo) {
    /* Explicit::updateTimes (Explicit.C:414-427) */
    if (J->t + J->dt > upTo) J->dt = upTo - J->t;
    J->inc++;
    predictor(J);
    if (!element_phases(J, J->dt)) return J->status;
    corrector(J);
    densities(J);
    /* a non-finite state stops the job in this cycle, before its end of
     * step (the reference has no such check and would run on) */
    if (!nodes_finite(J, 1 - J->f0)) {
      set_err(J, DNLG_E_NONFINITE, "non-finite nodal state");
      return J->status;
    }
    /* Explicit::endStep (Explicit.C:539-552): time, history rows, field swap */
    J->t += J->dt;
    write_history(J);
    J->f0 = 1 - J->f0;
    if (J->t < upTo) {
      if (!compute_jacobian(J, false)) return J->status;
      if (!compute_time_step(J, false)) return J->status;
    }
  }
  return DNLG_OK;
}

int oracle_job_sizes(const OracleJob *J, int64_t *n_nodes, int64_t *n_elems, int32_t *n_ip, int32_t *dim) {
  if (n_nodes) *n_nodes = J->N;
  if (n_elems) *n_elems = J->NE;
  if (n_ip) *n_ip = J->nip;
  if (dim) *dim = J->dim;
  return DNLG_OK;
}

int oracle_job_read_nodes(const OracleJob *J, double *coords, double *disp, double *speed) {
  if (coords) std::memcpy(coords, J->x.data(), 3 * J->N * sizeof(double));
  if (disp) std::memcpy(disp, J->disp.data(), 3 * J->N * sizeof(double));
  if (speed) std::memcpy(speed, J->fv[J->f0].data(), 3 * J->N * sizeof(double));
  return DNLG_OK;
}

int oracle_job_read_mass(const OracleJob *J, double *mass) {
  std::memcpy(mass, J->mass.data(), J->N * sizeof(double));
  return DNLG_OK;
}

int oracle_job_read_ip(const OracleJob *J, int field, double *out) {
  const int64_t G = J->NE * J->nip;
  auto scalar = [&](const std::vector<double> &v) {
    std::memcpy(out, v.data(), G * sizeof(double));
    return DNLG_OK;
  };
  auto tensor = [&](const std::vector<double> &v) {
    if (v.empty()) return DNLG_E_STATE;
    for (int c = 0; c < 6; c++)
      for (int64_t g = 0; g < G; g++) out[c * G + g] = v[g * 6 + c];
    return DNLG_OK;
  };
  switch (field) {
    case DNLG_F_PLASTIC_STRAIN: return scalar(J->ep);
    case DNLG_F_PLASTIC_STRAIN_RATE: return scalar(J->dep);
    case DNLG_F_GAMMA: return scalar(J->gam);
    case DNLG_F_GAMMA_CUMULATE: return scalar(J->gamc);
    case DNLG_F_YIELD_STRESS: return scalar(J->sy);
    case DNLG_F_TEMPERATURE: return scalar(J->T);
    case DNLG_F_DENSITY: return scalar(J->rho);
    case DNLG_F_PRESSURE:
      for (int64_t g = 0; g < G; g++) out[g] = third_trace(&J->sig[g * 6]);
      return DNLG_OK;
    case DNLG_F_INTERNAL_ENERGY: return scalar(J->eint);
    case DNLG_F_INELASTIC_ENERGY: return scalar(J->einel);
    case DNLG_F_VON_MISES:
      for (int64_t g = 0; g < G; g++) {
        const double *s = &J->sig[g * 6];
        out[g] = std::sqrt((sq(s[0] - s[3]) + sq(s[0] - s[5]) + sq(s[3] - s[5]) +
                            6 * (sq(s[1]) + sq(s[2]) + sq(s[4]))) / 2);
      }
      return DNLG_OK;
    case DNLG_F_STRESS: return tensor(J->sig);
    case DNLG_F_STRAIN: return tensor(J->eps);
    case DNLG_F_PLASTIC_STRAIN_TENSOR: return tensor(J->epsP);
  }
  return DNLG_E_INVALID;
}

/* Checkpoint / restart (dnlg_job_read_state / dnlg_job_write_state of the
 * engine, same record).  Writing into a job that has not been solved first
 * computes its reference Jacobian and masses from the mesh, as the pre-loop
 * does (Model.C:754-775, 537-590); the forced step and the element pass of
 * the pre-loop are what the state stands for.  The displacement
 * accumulator becomes coords - initial coordinates. */
int oracle_job_read_state(const OracleJob *J, dnlg_state *s) {
  const int64_t N = J->N, G = J->NE * J->nip;
  std::memcpy(s->coords, J->x.data(), 3 * N * sizeof(double));
  std::memcpy(s->disp_inc, J->fu[J->f0].data(), 3 * N * sizeof(double));
  std::memcpy(s->speed, J->fv[J->f0].data(), 3 * N * sizeof(double));
  std::memcpy(s->accel, J->fa[J->f0].data(), 3 * N * sizeof(double));
  for (int c = 0; c < 6; c++)
    for (int64_t g = 0; g < G; g++) s->stress[c * G + g] = J->sig[g * 6 + c];
  const std::vector<double> *rows[DNLG_STATE_IP_SCALARS] = {&J->ep, &J->dep, &J->gam, &J->gamc, &J->sy,
                                                             &J->T, &J->eint, &J->einel, &J->rho};
  for (int k = 0; k < DNLG_STATE_IP_SCALARS; k++)
    std::memcpy(s->ip_scalars + k * G, rows[k]->data(), G * sizeof(double));
  if (s->strain && !J->eps.empty())
    for (int c = 0; c < 6; c++)
      for (int64_t g = 0; g < G; g++) s->strain[c * G + g] = J->eps[g * 6 + c];
  if (s->plastic_strain_tensor && !J->epsP.empty())
    for (int c = 0; c < 6; c++)
      for (int64_t g = 0; g < G; g++) s->plastic_strain_tensor[c * G + g] = J->epsP[g * 6 + c];
  s->time = J->t;
  s->time_step = J->dt;
  s->increment = J->inc;
  return DNLG_OK;
}

int oracle_job_write_state(OracleJob *J, const dnlg_state *s) {
  if (J->status != DNLG_OK) return DNLG_E_STATE;
  if (J->prm.dt_method != DNLG_DT_COURANT || !(s->time_step > 0)) return DNLG_E_INVALID;
  if (!J->started) {
    J->started = true;
    if (!compute_jacobian(J, true)) return J->status;
    compute_mass(J);
  }
  const int64_t N = J->N, G = J->NE * J->nip;
  for (int64_t q = 0; q < 3 * N; q++) {
    const double x0 = J->x[q] - J->disp[q];
    J->x[q] = s->coords[q];
    J->disp[q] = s->coords[q] - x0;
  }
  for (int k = 0; k < 2; k++) {
    std::memcpy(J->fu[k].data(), s->disp_inc, 3 * N * sizeof(double));
    std::memcpy(J->fv[k].data(), s->speed, 3 * N * sizeof(double));
    std::memcpy(J->fa[k].data(), s->accel, 3 * N * sizeof(double));
  }
  for (int c = 0; c < 6; c++)
    for (int64_t g = 0; g < G; g++) J->sig[g * 6 + c] = s->stress[c * G + g];
  std::vector<double> *rows[DNLG_STATE_IP_SCALARS] = {&J->ep, &J->dep, &J->gam, &J->gamc, &J->sy,
                                                       &J->T, &J->eint, &J->einel, &J->rho};
  for (int k = 0; k < DNLG_STATE_IP_SCALARS; k++)
    std::memcpy(rows[k]->data(), s->ip_scalars + k * G, G * sizeof(double));
  if (!J->eps.empty() && s->strain)
    for (int c = 0; c < 6; c++)
      for (int64_t g = 0; g < G; g++) J->eps[g * 6 + c] = s->strain[c * G + g];
  if (!J->epsP.empty() && s->plastic_strain_tensor)
    for (int c = 0; c < 6; c++)
      for (int64_t g = 0; g < G; g++) J->epsP[g * 6 + c] = s->plastic_strain_tensor[c * G + g];
  J->t = s->time;
  J->dt = s->time_step;
  J->inc = s->increment;
  return DNLG_OK;
}

int oracle_job_scalars(const OracleJob *J, double *time, double *time_step, int64_t *increment) {
  if (time) *time = J->t;
  if (time_step) *time_step = J->dt;
  if (increment) *increment = J->inc;
  return DNLG_OK;
}

/* KE as Model::getTotalKineticEnergy (Model.C:612-625); IE and plastic work
 * integrated over the reference volume (no reference counterpart). */
int oracle_job_energies(const OracleJob *J, double out[4]) {
  double ke = 0.0;
  for (int64_t n = 0; n < J->N; n++) {
    double v2 = 0.0;
    for (int d = 0; d < 3; d++) v2 += sq(J->fv[J->f0][d * J->N + n]);
    ke += (J->mass.empty() ? 0.0 : J->mass[n]) * v2 / 2.0;
  }
  double ie = 0.0, wp = 0.0;
  for (int64_t g = 0; g < J->NE * J->nip; g++) {
    double vol = J->tab.w[g % J->nip] * J->detJ0[g];
    if (J->tab.axi) vol *= k2PI;
    ie += J->mat.density * vol * J->eint[g];
    wp += J->mat.density * vol * J->einel[g];
  }
  out[0] = ke;
  out[1] = ie;
  out[2] = wp;
  out[3] = ke + ie;
  return DNLG_OK;
}

int oracle_job_status(const OracleJob *J) { return J->status; }

/* HistoryFile::setSaveTime + add + Model::add (HistoryFile.C:131-280,
 * Model.C:927-939), rows only (no file); the engine's item conventions
 * (dnlg_history_item) */
int oracle_job_add_history(OracleJob *J, double start_time, double save_time, const dnlg_history_item *items,
                           int32_t n_items, int32_t *file_id) {
  if (J->started) return DNLG_E_STATE;
  for (int32_t k = 0; k < n_items; k++) {
    const dnlg_history_item &it = items[k];
    if (it.field < 0 || it.field >= DNLG_FIELD_COUNT) return DNLG_E_INVALID;
    if ((it.field >= DNLG_FIELD_STRAIN_INC && it.field < DNLG_FIELD_STRAIN_INC + 10) ||
        (it.field >= DNLG_FIELD_PLASTIC_STRAIN_INC && it.field < DNLG_FIELD_PLASTIC_STRAIN_INC + 10))
      return DNLG_E_INVALID;
    if (it.kind == DNLG_H_NODE && (it.index < 0 || it.index >= J->N)) return DNLG_E_INVALID;
    if (it.kind == DNLG_H_ELEMENT && (it.index < 0 || it.index >= J->NE || it.int_pt > J->nip)) return DNLG_E_INVALID;
  }
  if (J->node_elems.empty()) {
    /* Node::elements: the elements in creation (index) order */
    J->node_elems.resize(J->N);
    for (int64_t e = 0; e < J->NE; e++)
      for (int a = 0; a < J->nn; a++) J->node_elems[J->conn[e * J->nn + a]].push_back({e, a});
  }
  OracleJob::History h;
  h.start = start_time;
  h.next = start_time;
  h.save = save_time;
  h.items.assign(items, items + n_items);
  J->hist.push_back(h);
  if (file_id) *file_id = (int32_t)J->hist.size() - 1;
  return DNLG_OK;
}

int oracle_job_history_rows(const OracleJob *J, int32_t file_id, double *rows, int64_t max_rows, int64_t *n_rows) {
  if (file_id < 0 || file_id >= (int32_t)J->hist.size()) return DNLG_E_INVALID;
  const auto &h = J->hist[file_id];
  const int64_t width = 1 + (int64_t)h.items.size();
  const int64_t n = (int64_t)h.rows.size() / width;
  if (n_rows) *n_rows = n;
  if (rows && max_rows > 0) std::memcpy(rows, h.rows.data(), std::min(n, max_rows) * width * sizeof(double));
  return DNLG_OK;
}

/* Node::fieldScalar (Node.C:311-370) of every node, out [n_nodes] */
int oracle_job_read_node_field(OracleJob *J, int field, double *out) {
  if (field < 0 || field >= DNLG_FIELD_COUNT) return DNLG_E_INVALID;
  if (J->node_elems.empty()) {
    J->node_elems.resize(J->N);
    for (int64_t e = 0; e < J->NE; e++)
      for (int a = 0; a < J->nn; a++) J->node_elems[J->conn[e * J->nn + a]].push_back({e, a});
  }
  for (int64_t n = 0; n < J->N; n++) out[n] = node_value(J, field, n);
  return DNLG_OK;
}

/* the integrationPointsToNode table, [node][ip] */
int oracle_extrapolation(int type, double *E) {
  double e[8][8];
  extrapolation(type, e);
  const Tables T = make_tables(type);
  for (int a = 0; a < T.nn; a++)
    for (int p = 0; p < T.nip; p++) E[a * T.nip + p] = e[a][p];
  return DNLG_OK;
}

/* ---- unit-level entry points for pinning tests ---- */
void oracle_eig3(const double A[9], double Q[9], double w[3]) {
  double a[3][3], q[3][3];
  for (int i = 0; i < 9; i++) a[i / 3][i % 3] = A[i];
  kopp_eig3(a, q, w);
  for (int i = 0; i < 9; i++) Q[i] = q[i / 3][i % 3];
}

void oracle_polar_lnU(const double F[9], double dE[6], double R[9]) { polar_lnU(F, dE, R); }

void oracle_jacobi_eig3(const double A[9], double Q[9], double w[3]) {
  double a[3][3], q[3][3];
  for (int i = 0; i < 9; i++) a[i / 3][i % 3] = A[i];
  jacobi_eig3(a, q, w);
  for (int i = 0; i < 9; i++) Q[i] = q[i / 3][i % 3];
}

void oracle_rotate_sym(const double R[9], double S[6]) { rotate_sym(R, S); }

double oracle_law_yield(const dnlg_material *m, double epsp, double depsp, double T) {
  return law_yield(law_of(m), epsp, depsp, T);
}

double oracle_law_dyield(const dnlg_material *m, double epsp, double depsp, double T, double dtime) {
  return law_dyield(law_of(m), epsp, depsp, T, dtime);
}

/* one-IP radial return for unit tests: state[] = {sig[6], ep, dep, sy, gam,
 * gamc, T, eint, einel}; dE[6]; returns 0 or DNLG_E_NR_DIVERGED */
int oracle_radial_return(const dnlg_material *m, const double dE[6], double pressure, double timeStep,
                         double state[14]) {
  OracleJob J;
  J.nip = 1;
  J.mat = *m;
  J.law = law_of(m);
  J.sig.assign(state, state + 6);
  J.dE.assign(dE, dE + 6);
  J.pres.assign(1, pressure);
  J.ep.assign(1, state[6]);
  J.dep.assign(1, state[7]);
  J.sy.assign(1, state[8]);
  J.gam.assign(1, state[9]);
  J.gamc.assign(1, state[10]);
  J.T.assign(1, state[11]);
  J.eint.assign(1, state[12]);
  J.einel.assign(1, state[13]);
  if (!element_stress(&J, 0, timeStep)) return DNLG_E_NR_DIVERGED;
  for (int c = 0; c < 6; c++) state[c] = J.sig[c];
  state[6] = J.ep[0];
  state[7] = J.dep[0];
  state[8] = J.sy[0];
  state[9] = J.gam[0];
  state[10] = J.gamc[0];
  state[11] = J.T[0];
  state[12] = J.eint[0];
  state[13] = J.einel[0];
  return DNLG_OK;
}

/* shape tables: N [nip*nn], dNdr [nip*nn*3], w [nip]; returns nip */
int oracle_tables(int type, double *N, double *dNdr, double *w) {
  Tables T = make_tables(type);
  for (int p = 0; p < T.nip; p++) {
    w[p] = T.w[p];
    for (int a = 0; a < T.nn; a++) {
      N[p * T.nn + a] = T.N[p][a];
      for (int c = 0; c < 3; c++) dNdr[(p * T.nn + a) * 3 + c] = T.dNdr[p][a][c];
    }
  }
  return T.nip;
}

/* characteristic length of one element given its node coordinates
 * xyz[nn*3] (node-major) */
double oracle_char_length(int type, const double *xyz) {
  OracleJob J;
  J.tab = make_tables(type);
  J.dim = J.tab.dim;
  J.nn = J.tab.nn;
  J.N = J.nn;
  J.x.assign(3 * J.nn, 0.0);
  for (int a = 0; a < J.nn; a++)
    for (int c = 0; c < 3; c++) J.x[c * J.N + a] = xyz[a * 3 + c];
  J.conn.resize(J.nn);
  for (int a = 0; a < J.nn; a++) J.conn[a] = a;
  return char_length(&J, 0);
}

void oracle_chung_hulbert(double rho_b, double out[4]) {
  const double rb = rho_b;
  out[0] = (2.0 * rb - 1.0) / (1.0 + rb);
  out[1] = (5.0 - 3.0 * rb) / ((1.0 + rb) * (1.0 + rb) * (2.0 - rb));
  out[2] = 1.5 - out[0];
  out[3] = std::sqrt((12.0 * std::pow(1.0 + rb, 3) * (2.0 - rb)) /
                     (10.0 + 15.0 * rb - rb * rb + std::pow(rb, 3) - std::pow(rb, 4)));
}

double oracle_ramp_value(int type, double mn, double mx, double x) { return ramp_value(type, mn, mx, x); }

/* Elastic stiffness matrix of one element (node coordinates xyz[a * 3 + c]),
 * row-major (nn dim) x (nn dim), on its Jacobian (under != 0: the one-point
 * under-integrated matrix of PowerIterationUnder); 0, or
 * DNLG_E_NEG_JACOBIAN. */
int oracle_element_stiffness(int type, const double *xyz, double young, double poisson, int under, double *K) {
  OracleJob J;
  J.tab = make_tables(type);
  J.dim = J.tab.dim;
  J.nn = J.tab.nn;
  J.nip = J.tab.nip;
  J.N = J.nn;
  J.NE = 1;
  J.mat.young = young;
  J.mat.poisson = poisson;
  J.x.assign(3 * J.nn, 0.0);
  for (int a = 0; a < J.nn; a++)
    for (int c = 0; c < 3; c++) J.x[c * J.N + a] = xyz[a * 3 + c];
  J.conn.resize(J.nn);
  for (int a = 0; a < J.nn; a++) J.conn[a] = a;
  J.detJ.assign(J.nip, 0.0);
  J.detJ0.assign(J.nip, 0.0);
  J.radius.assign(J.nip, 0.0);
  J.dN.assign((size_t)J.nip * J.nn * J.dim, 0.0);
  if (under && J.nip > 1) return element_stiffness_under(&J, 0, K) ? DNLG_OK : DNLG_E_NEG_JACOBIAN;
  if (!element_jacobian(&J, 0, false)) return DNLG_E_NEG_JACOBIAN;
  element_stiffness(&J, 0, K);
  return DNLG_OK;
}

/* Power-iteration state of a job: the last maximum (omega_max^2) and the
 * iterations of the last time-step computation. */
int oracle_job_pi_state(const OracleJob *J, double *fmax, int *iterations) {
  if (fmax) *fmax = J->pi_fmax;
  if (iterations) *iterations = J->pi_iterations;
  return DNLG_OK;
}

}  // extern "C"
