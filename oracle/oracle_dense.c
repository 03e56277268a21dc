/*
 * oracle_dense.c -- dense FP64 kernels of the CPU oracle (test infrastructure
 * only, see oracle.h): GEMM, LU-based inverses and the compensated residual.
 *
 * These are tolerance stages: the reference calls LAPACK dgetrf/dgetri and
 * BLAS dgemm (ScaledDenseMatrix.cpp:232-279, :314-353) whose summation order
 * is implementation-defined, so nothing here is meant to be bit-compatible.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* C = A B, column-major n*n.  Four columns of C per task, blocked over k so
 * the A panel stays in cache; innermost loop is a contiguous axpy. */
void orc_gemm(int n, const double *A, const double *B, double *C) {
  const int JB = 4, KB = 128;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j0 = 0; j0 < n; j0 += JB) {
    int jn = MIN(JB, n - j0);
    double *c0 = C + (size_t)j0 * n;
    memset(c0, 0, sizeof(double) * (size_t)n * jn);
    for (int k0 = 0; k0 < n; k0 += KB) {
      int k1 = MIN(k0 + KB, n);
      for (int k = k0; k < k1; ++k) {
        const double *a = A + (size_t)k * n;
        if (jn == 4) {
          double b0 = B[k + (size_t)(j0 + 0) * n], b1 = B[k + (size_t)(j0 + 1) * n];
          double b2 = B[k + (size_t)(j0 + 2) * n], b3 = B[k + (size_t)(j0 + 3) * n];
          double *x0 = c0, *x1 = c0 + n, *x2 = c0 + 2 * (size_t)n, *x3 = c0 + 3 * (size_t)n;
          for (int i = 0; i < n; ++i) {
            double av = a[i];
            x0[i] = fma(av, b0, x0[i]); x1[i] = fma(av, b1, x1[i]);
            x2[i] = fma(av, b2, x2[i]); x3[i] = fma(av, b3, x3[i]);
          }
        } else {
          for (int q = 0; q < jn; ++q) {
            double b = B[k + (size_t)(j0 + q) * n];
            double *x = c0 + (size_t)q * n;
            for (int i = 0; i < n; ++i) x[i] = fma(a[i], b, x[i]);
          }
        }
      }
    }
  }
}

/* LU with partial pivoting (in place, column-major), then X = A^-1 by
 * solving A X = I column by column.  Returns ORC_ERR_SINGULAR on a zero pivot. */
int orc_inv_lu(int n, const double *Ain, double *X) {
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  int *piv = (int *)malloc(sizeof(int) * (size_t)n);
  if (!A || !piv) { free(A); free(piv); return ORC_ERR_OOM; }
  memcpy(A, Ain, sizeof(double) * (size_t)n * n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[k + (size_t)k * n]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(A[i + (size_t)k * n]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (best == 0.0) { free(A); free(piv); return ORC_ERR_SINGULAR; }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = A[k + (size_t)j * n];
        A[k + (size_t)j * n] = A[p + (size_t)j * n];
        A[p + (size_t)j * n] = t;
      }
    double inv = 1.0 / A[k + (size_t)k * n];
    for (int i = k + 1; i < n; ++i) A[i + (size_t)k * n] *= inv;
#pragma omp parallel for schedule(static) if (n - k > 256)
    for (int j = k + 1; j < n; ++j) {
      double u = A[k + (size_t)j * n];
      if (u == 0.0) continue;
      double *cj = A + (size_t)j * n;
      const double *ck = A + (size_t)k * n;
      for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
    }
  }
#pragma omp parallel for schedule(dynamic, 8)
  for (int j = 0; j < n; ++j) {
    double *x = X + (size_t)j * n;
    memset(x, 0, sizeof(double) * (size_t)n);
    x[j] = 1.0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
    for (int k = 0; k < n; ++k) { /* L y = P e_j (unit lower) */
      double xk = x[k];
      if (xk == 0.0) continue;
      const double *l = A + (size_t)k * n;
      for (int i = k + 1; i < n; ++i) x[i] -= l[i] * xk;
    }
    for (int k = n - 1; k >= 0; --k) { /* U x = y */
      const double *u = A + (size_t)k * n;
      x[k] /= u[k];
      double xk = x[k];
      for (int i = 0; i < k; ++i) x[i] -= u[i] * xk;
    }
  }
  free(A); free(piv);
  return ORC_OK;
}

/* SDM::inv (ScaledDenseMatrix.cpp:232-279): row 2-norms sL, column 2-norms sR
 * of the row-scaled matrix, LU inverse of diag(1/sL) A diag(1/sR), unscale
 * element (row j, col i) by (1/sL[i]) (1/sR[j]). */
int orc_inv_equilibrated(int n, const double *Ain, double *X) {
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *sL = (double *)calloc((size_t)n, sizeof(double));
  double *sR = (double *)calloc((size_t)n, sizeof(double));
  if (!A || !sL || !sR) { free(A); free(sL); free(sR); return ORC_ERR_OOM; }
  memcpy(A, Ain, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) sL[j] += A[(size_t)i * n + j] * A[(size_t)i * n + j];
  for (int j = 0; j < n; ++j) sL[j] = sqrt(sL[j]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[(size_t)i * n + j] *= 1 / sL[j];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) sR[i] += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    sR[i] = sqrt(sR[i]);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[(size_t)i * n + j] *= 1 / sR[i];
  int rc = orc_inv_lu(n, A, X);
  if (rc == ORC_OK) {
    for (int i = 0; i < n; ++i) sL[i] = 1 / sL[i];
    for (int i = 0; i < n; ++i) sR[i] = 1 / sR[i];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) X[(size_t)i * n + j] *= sL[i] * sR[j];
  }
  free(A); free(sL); free(sR);
  return rc;
}

/* R = I - A X.  Each entry is a compensated dot product (Ogita, Rump and
 * Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005,
 * algorithm Dot2): TwoProduct via fma, TwoSum on the running sum, errors
 * collected in a second accumulator, one final rounding. */
void orc_residual_dd(int n, const double *A, const double *X, double *R) {
  const int JB = 4;
#pragma omp parallel
  {
    double *s = (double *)malloc(sizeof(double) * (size_t)n * JB);
    double *c = (double *)malloc(sizeof(double) * (size_t)n * JB);
#pragma omp for schedule(dynamic, 1)
    for (int j0 = 0; j0 < n; j0 += JB) {
      int jn = MIN(JB, n - j0);
      for (int q = 0; q < jn; ++q)
        for (int i = 0; i < n; ++i) {
          s[(size_t)q * n + i] = (i == j0 + q) ? 1.0 : 0.0;
          c[(size_t)q * n + i] = 0.0;
        }
      for (int k = 0; k < n; ++k) {
        const double *a = A + (size_t)k * n;
        for (int q = 0; q < jn; ++q) {
          double x = X[k + (size_t)(j0 + q) * n];
          double *sq = s + (size_t)q * n, *cq = c + (size_t)q * n;
          for (int i = 0; i < n; ++i) {
            double P = a[i] * x;
            double E = fma(a[i], x, -P); /* a*x = P + E exactly */
            double p = -P;
            double t = sq[i] + p;        /* TwoSum(s, p) */
            double bb = t - sq[i];
            double err = (sq[i] - (t - bb)) + (p - bb);
            sq[i] = t;
            cq[i] = cq[i] + (err - E);
          }
        }
      }
      for (int q = 0; q < jn; ++q)
        for (int i = 0; i < n; ++i)
          R[i + (size_t)(j0 + q) * n] = s[(size_t)q * n + i] + c[(size_t)q * n + i];
    }
    free(s);
    free(c);
  }
}
