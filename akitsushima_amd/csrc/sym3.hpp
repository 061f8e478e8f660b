// Small 3x3 matrix helpers for the flexible-cell barostat: row-major
// products, the inverse, and the eigen-decomposition of a symmetric matrix
// by cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th
// ed., sec. 8.5), which a single device thread runs in a few sweeps.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ak {

// C = A B (row-major 3x3); C must not alias A or B.
__host__ __device__ inline void mat3_mul(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// det(A) and, into B, the inverse by the adjugate.
__host__ __device__ inline double mat3_inv(const double* A, double* B) {
  const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8],
               c02 = A[3] * A[7] - A[4] * A[6];
  const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
  const double r = 1.0 / det;
  B[0] = c00 * r;
  B[1] = (A[2] * A[7] - A[1] * A[8]) * r;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * r;
  B[3] = c01 * r;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * r;
  B[5] = (A[2] * A[3] - A[0] * A[5]) * r;
  B[6] = c02 * r;
  B[7] = (A[1] * A[6] - A[0] * A[7]) * r;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * r;
  return det;
}

// Symmetric A = U diag(w) U^T; the columns of U (row-major) are the
// eigenvectors.  Sweeps until the off-diagonal part is below 1e-34 of the
// squared norm (machine precision) or is exactly zero.
__host__ __device__ inline void sym3_eig(const double* A, double* w, double* U) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = 0.5 * (A[3 * i + j] + A[3 * j + i]);
      U[3 * i + j] = i == j ? 1.0 : 0.0;
    }
  const double nrm = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2] +
                     2.0 * (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2]);
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off == 0.0 || off <= 1e-34 * nrm) break;
    for (int r = 0; r < 3; ++r) {
      const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      // J = [[c, s], [-s, c]] in the (p, q) plane zeroes a'_pq of J^T a J
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - s * akq;
        a[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - s * aqk;
        a[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; ++k) {
        const double ukp = U[3 * k + p], ukq = U[3 * k + q];
        U[3 * k + p] = c * ukp - s * ukq;
        U[3 * k + q] = s * ukp + c * ukq;
      }
    }
  }
  for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// B = U diag(f) U^T.
__host__ __device__ inline void sym3_compose(const double* U, const double* f, double* B) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      B[3 * i + j] = U[3 * i] * f[0] * U[3 * j] + U[3 * i + 1] * f[1] * U[3 * j + 1] +
                     U[3 * i + 2] * f[2] * U[3 * j + 2];
}

}  // namespace ak
