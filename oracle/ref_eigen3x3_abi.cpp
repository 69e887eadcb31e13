/*
 * C-linkage shim around the reference's own symmetric 3x3 eigensolver
 * (Sources/dnlMaths/Eigen3x3.C:454-634, dsyevd3), compiled from the
 * reference source where it lies into oracle/_ref/.  TEST INFRASTRUCTURE:
 * used only by tests/ to pin the oracle's restatement bit-for-bit.
 */
int dsyevd3(double A[3][3], double Q[3][3], double w[3]);

extern "C" int ref_dsyevd3(const double *A9, double *Q9, double *w3) {
  double A[3][3], Q[3][3];
  for (int i = 0; i < 9; i++) A[i / 3][i % 3] = A9[i];
  const int rc = dsyevd3(A, Q, w3);
  for (int i = 0; i < 9; i++) Q9[i] = Q[i / 3][i % 3];
  return rc;
}
