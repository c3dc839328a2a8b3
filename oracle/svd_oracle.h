/*
 * svd_oracle.h -- CPU restatement of the ddc-svd host path, used ONLY as a
 * test oracle and as the host-core baseline timed beside the GPU path.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  The product library (libddcsvd)
 * never links or calls it.
 *
 * Every function restates one reference routine; the citation is given at
 * each declaration.  Storage conventions are the reference's: column-major
 * matrices, unit-norm Householder vectors (H = I - 2 v v^T) stored in place,
 * singular values in ascending order.
 */
#ifndef DDCSVD_ORACLE_H
#define DDCSVD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Golub-Kahan bidiagonalization, serial.  bidiag.c:33-186 (with
 * left_householder bidiag.c:188-218, right_householder bidiag.c:220-250).
 * A (m x n, column-major) is overwritten with the reflectors; alpha gets
 * min(m,n) entries, beta min(m,n)-1 (m >= n) or m (m < n). */
void orc_bidiag_seq(int m, int n, double *A, double *alpha, double *beta);

/* The first `steps` steps of orc_bidiag_seq's main loop (bidiag.c:73-131)
 * and nothing else: a bounded sample of the serial reduction for timing. */
void orc_bidiag_sample(int m, int n, double *A, double *alpha, double *beta,
                       int steps);

/* Singular values of the N x (N+1) upper bidiagonal [b1; b2] by the
 * reference's double divide and conquer.  Calculations-Parallel.c:852-874
 * and the routines it calls (:48-348, :351-588, :590-703, :706-850).
 * b1, b2 have N entries each; sigma gets N values in ascending order. */
void orc_dc_values(int N, const double *b1, const double *b2, double *sigma);

/* Right singular vectors of the n x m bidiagonal (m = n or n+1) by the
 * squared-matrix twisted factorization.  parallel-twisted.c:554-637.
 * X is n rows of m entries (row i = vector of sigma[i]). */
void orc_right_vectors(int n, int m, const double *a, const double *b,
                       const double *sigma, double *X);

/* Y_i = B X_i / sigma_i.  parallel-twisted.c:530-551 with BidiagMatVec :58-86.
 * X rows have stride m, Y rows have stride n. */
void orc_left_vectors(int n, int m, const double *a, const double *b,
                      const double *sigma, const double *X, double *Y);

/* Back-transformation of one vector.  multU bidiag_par.c:1046-1095,
 * multV bidiag_par.c:990-1043.  fix_multv = 0 keeps the reference's
 * indexing (leading dimension m at :1034,1036 and X stride min(m,n) at
 * :1025, both wrong when m != n); fix_multv = 1 uses leading dimension n
 * and X stride len_beta+1. */
void orc_mult_u(int m, int n, int vec, const double *A_mod, const double *Y,
                double *U);
void orc_mult_v(int m, int n, int vec, const double *AT, const double *X,
                double *V, int fix_multv);

/* Whole host path with the serial bidiagonalization in place of the OpenCL
 * one, orchestrated as svd_gpu.c:53-131.  U is m x m (first min(m,n)
 * columns written), V is n x n (first min(m,n) columns written).  Optional
 * outputs alpha/beta (may be NULL).  Returns 0 on success, -1 on allocation
 * failure. */
int orc_svd(int m, int n, double *A, double *sigma, double *U, double *V,
            double *alpha_out, double *beta_out, int fix_multv);

/* Stage timings of the last orc_svd call in seconds:
 * [bidiag, values, vectors, back-transformation]. */
void orc_last_timings(double out[4]);

/* Explicit orthogonal factors of the bidiagonalization from the stored
 * reflectors: U (m x m) as form_u (bidiag.c:252-302, = form_u_par
 * bidiag_par.c:877-932) and V (n x n) as form_v (bidiag.c:305-359, =
 * form_v_par bidiag_par.c:934-988).  A_mod is m x n, leading dimension m. */
void orc_form_u(int m, int n, const double *A_mod, double *U);
void orc_form_v(int m, int n, const double *A_mod, double *V);

#ifdef __cplusplus
}
#endif
#endif
