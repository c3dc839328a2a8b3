/*
 * ref_driver.c -- entry points around the reference's own host routines,
 * compiled together with the reference sources (where they lie) into
 * oracle/_ref/libddcref.so.  TEST INFRASTRUCTURE ONLY: used to pin the
 * oracle restatement and as the "reference" CPU baseline.
 *
 * ref_svd_seq mirrors the orchestration of svd_gpu.c:53-131 with the
 * serial bidiagonalization bidiag_seq (bidiag.c:33) in place of the OpenCL
 * bidiag_par, since no OpenCL device exists on these hosts.  beta gets one
 * spare zero entry because the value solver reads beta[min(m,n)-1]
 * (Calculations-Parallel.c:606,627).
 */
#include <stdlib.h>
#include <string.h>

#include "Calculations-Parallel.h"
#include "bidiag.h"
#include "cl-helper.h"
#include "bidiag_par.h"
#include "matrix_helper.h"
#include "parallel-twisted.h"

void ref_bidiag_seq(int m, int n, double *A, double *alpha, double *beta) {
  bidiag_seq(m, n, A, alpha, beta);
}

void ref_dc_values(int N, double *b1, double *b2, double *sigma) {
  GetSingularValues_Parallel(N, b1, b2, sigma);
}

void ref_right_vectors(int n, int m, double *a, double *b, double *sigma,
                       double *X) {
  CalcRightSingularVectors(n, m, a, b, sigma, X);
}

void ref_left_vectors(int n, int m, double *a, double *b, double *sigma,
                      double *X, double *Y) {
  RighttoLeftSingularVectors(n, m, a, b, sigma, X, Y);
}

int ref_svd_seq(int m, int n, double *A, double *sigma, double *U, double *V,
                double *alpha_out, double *beta_out) {
  int mn = (m >= n) ? n : m;
  int len_beta = (m >= n) ? n - 1 : m;
  double *AT = (double *)malloc(sizeof(double) * (size_t)m * n);
  double *alpha = (double *)malloc(sizeof(double) * mn);
  double *beta = (double *)calloc((size_t)len_beta + 1, sizeof(double));
  double *X = (double *)malloc(sizeof(double) * (size_t)mn * (len_beta + 1));
  double *Y = (double *)malloc(sizeof(double) * (size_t)mn * mn);
  if (!AT || !alpha || !beta || !X || !Y) return -1;
  bidiag_seq(m, n, A, alpha, beta);
  transpose(m, n, A, AT);
  GetSingularValues_Parallel(mn, alpha, beta, sigma);
  CalcRightSingularVectors(mn, len_beta + 1, alpha, beta, sigma, X);
  RighttoLeftSingularVectors(mn, len_beta + 1, alpha, beta, sigma, X, Y);
#pragma omp parallel for
  for (int i = 0; i < mn; i++) {
    multU(m, n, i, A, Y, U + (size_t)i * m);
    multV(m, n, i, AT, X, V + (size_t)i * n);
  }
  if (alpha_out) memcpy(alpha_out, alpha, sizeof(double) * mn);
  if (beta_out) memcpy(beta_out, beta, sizeof(double) * len_beta);
  free(AT); free(alpha); free(beta); free(X); free(Y);
  return 0;
}

/* glibc rand() stream exactly as test-whole-svd.c:18-24,69-73 draws it:
 * entries lo + (rand()/(RAND_MAX+1.0))*(hi-lo), after srand(seed). */
void ref_fill_rand(int count, double lo, double hi, unsigned seed, double *out) {
  srand(seed);
  for (int i = 0; i < count; ++i) {
    double c = rand() / (1.0 * RAND_MAX + 1);
    out[i] = lo + c * (hi - lo);
  }
}
