/*
 * svd_gpu.h -- C interface of libddcsvd, the MI355X (gfx950) implementation
 * of the ddc-svd SVD path.
 *
 * Link with -lddcsvd (plus the HIP runtime, which the library pulls in).
 * All matrices are column-major doubles.  Singular values come out in
 * ascending order and column i of U and V pairs with sigma[i], as in the
 * reference (Calculations-Parallel.c:56, svd_gpu.c:117-121).
 */
#ifndef DDCSVD_SVD_GPU_H
#define DDCSVD_SVD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the int-returning entry points. */
#define SVD_GPU_OK 0
#define SVD_GPU_E_ARG -1     /* invalid dimensions or pointers */
#define SVD_GPU_E_NOMEM -2   /* device or host allocation failed */
#define SVD_GPU_E_HIP -3     /* a HIP runtime call failed */
#define SVD_GPU_E_NODEV -4   /* no usable GPU */
#define SVD_GPU_E_NUMERIC -5 /* non-finite input */

/*
 * Drop-in replacement of the reference entry point
 *   void svd_gpu(int m, int n, double* A, double* sigma, double* U, double* V)
 * declared at svd_gpu.h:5 and defined at svd_gpu.c:53-131.
 *   A      m x n host buffer; overwritten (as in the reference, with the
 *          Householder reflectors of the bidiagonalization for m >= n)
 *   sigma  min(m,n) singular values, ascending
 *   U      m x m host buffer; the first min(m,n) columns are written
 *   V      n x n host buffer; the first min(m,n) columns are written
 *          (V itself, not V^T)
 * Errors print a message to stderr and abort(), as the reference does
 * (cl-helper.h:47-55, svd_gpu.c:85-96).  Thread-safe (calls serialise).
 */
void svd_gpu(int m, int n, double *A, double *sigma, double *U, double *V);

/* svd_gpu with the same arguments and results, returning a status code
 * instead of aborting.  No reference counterpart (the reference has only the
 * aborting form, svd_gpu.c:53). */
int svd_gpu_status(int m, int n, double *A, double *sigma, double *U,
                   double *V);

/*
 * Device-pointer form of svd_gpu for callers whose data already lives in HBM
 * (no reference counterpart: the reference copies host buffers in and out on
 * every call, bidiag_par.c:298-301,405-418).
 *   dA     m x n device matrix, leading dimension lda >= m; NOT modified
 *   dsigma min(m,n) device values, ascending
 *   dU     m x min(m,n) device matrix, leading dimension ldu >= m
 *   dV     n x min(m,n) device matrix, leading dimension ldv >= n
 *   stream a hipStream_t (NULL = the null stream, so the call is also
 *          ordered after the caller's work on blocking streams); the call
 *          is ordered on it and returns after the results are complete.
 */
int svd_gpu_dev(int m, int n, const double *dA, int lda, double *dsigma,
                double *dU, int ldu, double *dV, int ldv, void *stream);

/*
 * Stand-alone bidiagonalization, the replacement of
 *   void bidiag_par(int m, int n, double* A, double* alpha, double* beta)
 * (bidiag_par.h:30, bidiag_par.c:34-450).  Host buffers.  A is overwritten
 * with the unit-norm reflectors exactly as the reference stores them
 * (columns below/on the diagonal: left, rows right of the diagonal: right);
 * alpha gets min(m,n) entries.  m >= n: beta gets n-1 entries (n x n upper
 * bidiagonal).  m < n: beta gets m entries, the m x (m+1) upper bidiagonal of
 * the reference (bidiag.c:133-160, bidiag_par.c:361-383): the last row takes
 * a right reflector over columns m..n-1 and no left one, alpha[m-1] is taken
 * as it stands and A[m-1, m-1] is set to zero.
 */
int bidiag_gpu(int m, int n, double *A, double *alpha, double *beta);

/*
 * Explicit orthogonal factors of the bidiagonalization A = U B V^T, the
 * replacements of form_u_par / form_v_par (bidiag_par.h:33-34,
 * bidiag_par.c:877-988, equal to form_u / form_v of bidiag.c:252-359).
 * A (m x n, m >= n, host) holds the reflectors as bidiag_gpu leaves them;
 * U gets m x m, V gets n x n (host, column-major).
 */
int bidiag_form_u_gpu(int m, int n, const double *A, double *U);
int bidiag_form_v_gpu(int m, int n, const double *A, double *V);

/*
 * Singular values (ascending) of the n x n upper bidiagonal with diagonal
 * alpha[n] and super-diagonal beta[n-1], by the GPU divide and conquer
 * (replacement of GetSingularValues_Parallel, Calculations-Parallel.h:41,
 * Calculations-Parallel.c:852).  refine != 0 also runs the per-value
 * twisted-factorization refinement.  Host buffers.
 */
int bdsvd_values_gpu(int n, const double *alpha, const double *beta,
                     double *sigma, int refine);

/*
 * Singular vectors of the n x n upper bidiagonal for its ascending singular
 * values (replacement of CalcRightSingularVectors and
 * RighttoLeftSingularVectors, parallel-twisted.h:18-19).  X and Y are n x n
 * host matrices whose column i is the right / left vector of sigma[i];
 * sigma is refined in place.  Host buffers.
 */
int bdsvd_vectors_gpu(int n, const double *alpha, const double *beta,
                      double *sigma, double *X, double *Y);

/*
 * Reverse-mode derivative of the thin SVD (no reference counterpart: the
 * reference has no derivative).  Given A = U diag(sigma) V^T as returned by
 * svd_gpu_dev (U m x min(m,n), V n x min(m,n), sigma ascending) and the
 * cotangents gU, gsigma, gV of a scalar loss (each may be NULL = zero),
 * writes gA = dL/dA (m x n).  Device pointers, column-major.
 *   gap_guard  regularises 1/(sigma_j^2 - sigma_i^2) for near-equal values:
 *              below delta = gap_guard * sigma_max^2 the factor is d/delta^2
 *              (1e-13 is a good default; 0 = unguarded)
 * The formula is exact for distinct nonzero singular values.
 */
int svd_gpu_backward(int m, int n, const double *U, int ldu, const double *sigma,
                     const double *V, int ldv, const double *gU, int ldgu,
                     const double *gsigma, const double *gV, int ldgv, double *gA,
                     int ldga, double gap_guard, void *stream);

/*
 * Mixed-precision SVD of a square n x n FP64 matrix (no reference
 * counterpart: the reference computes in FP64 only, svd_gpu.c:53).  The
 * bidiagonalization stores the matrix and its reflectors in FP32 (vector
 * arithmetic FP64) and runs its blocked trailing update on the matrix cores
 * with FP32 accumulation in the chosen precision; the bidiagonal solve is
 * FP64 and the back-transformation FP32 on the matrix cores; then
 * refine_steps (0..8) refinement steps (Ogita-Aishima: U^T A V, U^T U and
 * V^T V in FP64, the update products U F and V G in FP32) restore accuracy
 * where the FP32 stage resolved the singular vectors.  Device pointers as in svd_gpu_dev; results ascending,
 * A not modified.
 *   update       SVD_GPU_MIXED_F32, _BF16X3 (BF16 products on the two-term
 *                BF16 split: hi*hi + hi*lo + lo*hi) or _BF16 (one product),
 *                optionally ORed with SVD_GPU_MIXED_SEPARATED: first-order
 *                corrections up to 1e-2 instead of 1e-3.  For spectra whose
 *                close values come in isolated pairs (e.g. Gaussian
 *                matrices) two steps then reach FP64 accuracy; spectra with
 *                long runs of values closer than the FP32 stage's error
 *                (e.g. geometrically graded ones) lose orthogonality with it;
 *                and/or with SVD_GPU_MIXED_RESOLVE (needs refine_steps >= 1):
 *                after the last step, the values of the pairs it left
 *                unresolved (gaps below the FP32 stage's error) get a
 *                Rayleigh-Ritz step in FP64, A_R = U_R^T A V_R = P S Q^T,
 *                U_R <- U_R P, V_R <- V_R Q, with an FP64 SVD of the block
 *                A_R
 *   diag         optional host array of 8 (10 with SVD_GPU_MIXED_RESOLVE):
 *                [0] ||I-U^T U||_F, [1] ||I-V^T V||_F before refinement,
 *                [2] bidiagonalization, [3] values and vectors,
 *                [4] back-transformation, [5] refinement (resolve
 *                included), [6] FP32
 *                conversion (ms), [7] max |sigma change| / sigma_max in the
 *                last refinement step, [8] the number of values resolved,
 *                [9] the resolve stage (ms)
 * Synchronous: returns after the results are complete.
 */
#define SVD_GPU_MIXED_F32 1
#define SVD_GPU_MIXED_BF16X3 2
#define SVD_GPU_MIXED_BF16 3
#define SVD_GPU_MIXED_SEPARATED 0x100
#define SVD_GPU_MIXED_RESOLVE 0x200
int svd_gpu_mixed(int n, const double *dA, int lda, double *dsigma, double *dU,
                  int ldu, double *dV, int ldv, int update, int refine_steps,
                  double *diag, void *stream);

/*
 * Batched SVD of many small matrices of uneven shapes (no reference
 * counterpart: the reference factors one matrix per call, svd_gpu.c:53).
 * One persistent launch, replayed from a captured HIP graph, factors every
 * matrix with one-sided Jacobi; 1 <= m, n <= 256 per matrix.
 *
 * Buffers are packed back to back in batch order, column-major, on the
 * device; svd_gpu_batched_offsets gives the element offsets (count + 1
 * entries each, the last one the total):
 *   A_k      m_k x n_k        at dA + offA[k]
 *   sigma_k  min(m_k, n_k)    at dsigma + offS[k], ascending
 *   U_k      m_k x min(m_k,n_k) at dU + offU[k]
 *   V_k      n_k x min(m_k,n_k) at dV + offV[k]
 * dtype selects double (SVD_GPU_F64) or float (SVD_GPU_F32) buffers.
 * A is not modified.  Columns of U (or of V for a wide matrix) that belong
 * to an exactly zero singular value are completed to an orthonormal set.
 */
#define SVD_GPU_F64 0
#define SVD_GPU_F32 1
typedef struct svd_gpu_batch svd_gpu_batch;

/* Describes the batch; allocates the device descriptors and workspace. */
int svd_gpu_batched_plan(svd_gpu_batch **plan, int count, const int *m,
                         const int *n, int dtype);
/* Element offsets of the packed buffers (any pointer may be NULL). */
int svd_gpu_batched_offsets(const svd_gpu_batch *plan, long long *offA,
                            long long *offS, long long *offU, long long *offV);
/*
 * Factors the batch, ordered on stream (NULL = the null stream); returns once
 * the work is enqueued.  dinfo (device, count ints, may be NULL) receives the
 * Jacobi sweeps each matrix took, or -1 if it did not converge.  The first
 * call autotunes the launch (if svd_gpu_batched_tune was not called) and
 * captures a graph; later calls with the same buffers replay it and allocate
 * nothing.  Calls on one plan must not overlap.
 */
int svd_gpu_batched_run(svd_gpu_batch *plan, const void *dA, void *dsigma,
                        void *dU, void *dV, int *dinfo, void *stream);
/* Times the kernel variants, launch sizes and cooperation thresholds on the
 * whole batch (the result is also remembered for later plans of similar
 * batches); writes the outputs.  Synchronous. */
int svd_gpu_batched_tune(svd_gpu_batch *plan, const void *dA, void *dsigma,
                         void *dU, void *dV, void *stream);
/* Chosen configuration: kernel variant, resident 4-wave workgroups per CU,
 * the smallest min(m, n) that a whole workgroup factors together, and the
 * column-block widths for matrices of up to 128 / up to 256 rows.  Returns
 * the number of variants (variant = -1 before tuning). */
int svd_gpu_batched_config(const svd_gpu_batch *plan, int *variant,
                           int *wgs_per_cu, int *coop_min_n, int *block_small,
                           int *block_large);
/* Forces a configuration (variant < number of variants, 1..8 workgroups per
 * CU, coop_min_n >= 1; 257 disables cooperation). */
int svd_gpu_batched_set_config(svd_gpu_batch *plan, int variant, int wgs_per_cu,
                               int coop_min_n);
void svd_gpu_batched_destroy(svd_gpu_batch *plan);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the batched SVD without a plan, with the same packed
 * buffer layout as svd_gpu_batched_offsets and the same kernel as
 * svd_gpu_batched_run, at a fixed launch configuration (variant 1, two
 * workgroups per CU, cooperation from min(m, n) = 32).  Its descriptors,
 * queue head and working copies live in a 32 MiB device array that each GPU
 * holds from the library's first kernel launch until the process ends; a
 * batch needing more (about sum of max(m,n) min(m,n) + min(m,n)^2 elements,
 * 8 or 4 bytes each), or more than 4096 matrices, is rejected with
 * SVD_GPU_E_ARG.  Calls no hipMalloc/hipFree; copies the descriptors from
 * host arrays and synchronizes the stream once before the launch.  Calls
 * share the array: calls that overlap on one GPU race.
 */
int svd_gpu_batched_noalloc(int count, const int *m, const int *n, int dtype,
                            const void *dA, void *dsigma, void *dU, void *dV,
                            int *dinfo, void *stream);

/*
 * The library's FP64 GEMM on the matrix cores (no reference counterpart; the
 * kernel of the blocked trailing update and of the back-transformation),
 * exposed for callers and tests: C = alpha op(A) op(B) + beta C on device
 * column-major matrices, op(X) = X for 'N', X^T for 'T'; op(A) is M x K,
 * op(B) K x N.  Ordered on stream (NULL = the null stream); returns once
 * enqueued.
 */
int svd_gpu_dgemm(char transa, char transb, int M, int N, int K, double alpha,
                  const double *dA, int lda, const double *dB, int ldb, double beta,
                  double *dC, int ldc, void *stream);

/* Milliseconds of the stages of the last successful svd_gpu / svd_gpu_status
 * / svd_gpu_dev call in this process: [0] copy-in, [1] bidiagonalization,
 * [2] singular values, [3] vectors, [4] back-transformation, [5] copy-out.
 * With kernel profiling on (svd_gpu_profile_kernels): [6] / [7] summed
 * durations of the left / right GEMV launches of the bidiagonalization and
 * [8] / [9] their launch counts, [10] / [11] those of the one-pass GEMV
 * (both products of a column step in one read of the trailing matrix)
 * (zeros otherwise).  Returns the number of entries written (at most
 * count, at most 12). */
int svd_gpu_last_timings(double *ms, int count);

/* Measurement hook (no reference counterpart): on != 0 brackets every GEMV
 * launch of the bidiagonalization with HIP events on its stream, reported by
 * svd_gpu_last_timings; costs two event records per launch. */
int svd_gpu_profile_kernels(int on);

/* Measurement hook (no reference counterpart): the first `steps` left-GEMV
 * launches (1 <= steps <= n-1) of the bidiagonalization of a copy of the
 * m x n device matrix dA (m >= n >= 2), with no other kernel of those column
 * steps in between, for hardware-counter passes over that kernel alone.  The
 * launches read and write only library workspace; results are discarded. */
int svd_gpu_profile_left_gemv(int m, int n, const double *dA, int lda, int steps,
                              void *stream);

/* Measurement hook (no reference counterpart): the first `steps` right-GEMV
 * launches (1 <= steps <= n-2) of the bidiagonalization of a copy of the
 * m x n device matrix dA (m >= n >= 3), after one left GEMV of column 0 and
 * with no other kernel in between: each launch streams the trailing block of
 * its column step, as in the SVD.  Workspace only; results are discarded. */
int svd_gpu_profile_right_gemv(int m, int n, const double *dA, int lda, int steps,
                               void *stream);

/* Measurement hook (no reference counterpart): the first `steps` launches
 * (1 <= steps <= n-2) of the one-pass GEMV (both products of a column step
 * in one read of the trailing matrix) of the bidiagonalization of a copy of
 * the m x n device matrix dA (8192 >= m >= n >= 3), with no other kernel of
 * those column steps in between, for hardware-counter passes over that
 * kernel alone.  The launches read and write only library workspace;
 * results are discarded. */
int svd_gpu_profile_fused_gemv(int m, int n, const double *dA, int lda, int steps,
                               void *stream);

/* Human-readable description of the last failure (empty if none). */
const char *svd_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
