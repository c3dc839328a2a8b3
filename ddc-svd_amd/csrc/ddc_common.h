// ddc_common.h -- shared internals of libddcsvd (gfx950, HIP).
//
// Error model: every internal routine returns an int status (0 = success,
// negative = DDC_E_* code).  The reference-compatible svd_gpu() turns a
// failure into a message on stderr and abort(), as the reference does
// (cl-helper.h:47-55, svd_gpu.c:85-96); svd_gpu_status() returns the code.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>

#define DDC_OK 0
#define DDC_E_ARG -1
#define DDC_E_NOMEM -2
#define DDC_E_HIP -3
#define DDC_E_NODEV -4
#define DDC_E_NUMERIC -5

namespace ddc {

// Records the first HIP failure of the calling thread so entry points can
// report it; returns DDC_E_HIP on failure.
int hip_fail(hipError_t e, const char* what, const char* file, int line);
// Optional HIP-event timing of single kernel launches (for measurement; off
// unless svd_gpu_profile_kernels(1) was called).  kclock_on() tells whether
// it is active; kclock_mark() records an event on st and returns its index;
// kclock_span(cls, a, b) books the span between two marks to class cls.
// kclock_pair() hands out two events for hipExtLaunchKernel, which stamps
// them with the start and the end of that one kernel, and books the span.
bool kclock_on();
int kclock_mark(hipStream_t st);
void kclock_span(int cls, int a, int b);
int kclock_pair(int cls, hipEvent_t* e0, hipEvent_t* e1);
enum { KCLOCK_LEFT_GEMV = 0, KCLOCK_RIGHT_GEMV = 1, KCLOCK_FUSED_GEMV = 2, KCLOCK_CLASSES = 3 };

// Message returned by svd_gpu_last_error(); clear_error() empties it.
void report_error(const char* msg);
void clear_error();

#define DDC_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return ::ddc::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define DDC_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != 0) return rc_;  \
  } while (0)

// Kernel launches are checked through hipGetLastError after the launch.
#define DDC_LAUNCHED() DDC_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// A bump allocator over one device allocation; every stage asks it for the
// buffers it needs, so one svd_gpu call makes a single hipMalloc.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    size_t bytes = round_up((int64_t)(count * sizeof(T)), 256);
    if (used + bytes > cap) return nullptr;
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

// Bytes an Arena needs for take<T>(count) calls, to size the allocation.
struct ArenaPlan {
  size_t bytes = 0;
  template <typename T>
  void add(size_t count) { bytes += round_up((int64_t)(count * sizeof(T)), 256); }
};

// ---------------------------------------------------------------------------
// Stage entry points (device pointers, stream-ordered).  Each has a
// *_plan() that adds its workspace to an ArenaPlan.
// ---------------------------------------------------------------------------

// Golub-Kahan bidiagonalization of the m x n (m >= n) column-major matrix A
// (leading dimension lda, rows m..lda-1 must hold zeros) in place: reflectors
// as the reference stores them (bidiag_par.c:34-450), alpha[n], beta[n] with
// beta[n-1] = 0.  nb is the panel width of the blocked update.
void bidiag_plan(ArenaPlan& p, int m, int n, int lda);
// Measurement: `steps` left-GEMV launches alone (A and alpha overwritten).
int bidiag_left_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps, Arena& ws,
                          hipStream_t st);
// Measurement: `steps` one-pass GEMV launches alone (m <= 8192).
int bidiag_fused_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps,
                           Arena& ws, hipStream_t st);
// Measurement: `steps` right-GEMV launches (after one left GEMV at k = 0).
int bidiag_right_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps,
                           Arena& ws, hipStream_t st);
int bidiag_run(int m, int n, double* A, int lda, double* alpha, double* beta,
               Arena& ws, hipStream_t st);

// The same with the matrix and its reflectors stored in FP32 (vector
// arithmetic stays FP64) and the blocked trailing update in reduced
// precision on the matrix cores: UPD_F32 (FP32 GEMM), UPD_BF16X3 (BF16
// GEMM on the two-term BF16 split, three products), UPD_BF16 (one BF16
// product); FP32 accumulation in every case.
enum { UPD_F32 = 1, UPD_BF16X3 = 2, UPD_BF16 = 3 };
void bidiag_plan_f32(ArenaPlan& p, int m, int n, int lda, int upd);
int bidiag_run_f32(int m, int n, float* A, int lda, double* alpha, double* beta, int upd,
                   Arena& ws, hipStream_t st);

// Mixed-precision SVD of the n x n FP64 matrix A (device pointers): FP32
// bidiagonalization with update mode upd, FP64 bidiagonal solve, FP32
// back-transformation, `steps` FP64 refinement steps (separated: the larger
// bound on the first-order corrections; resolve: an FP64 Rayleigh-Ritz step
// on the values whose pairs the last step left unresolved).  diag: see
// mixed.hip.
void mixed_plan(ArenaPlan& p, int n, int upd, bool resolve);
int mixed_run(int n, const double* A, int lda, double* sigma, double* U, int ldu, double* V,
              int ldv, int upd, bool separated, bool resolve, int steps, double* diag,
              Arena& ws, hipStream_t st);

// FP64 SVD of the m x n (m >= n) device matrix dA on the caller's arena and
// stream (the plain pipeline; capi.hip): sigma ascending, U m x n, V n x n.
void svd_device_plan(ArenaPlan& p, int m, int n);
int svd_device(int m, int n, const double* dA, int lda, double* dsigma, double* dU, int ldu,
               double* dV, int ldv, Arena& ws, hipStream_t st);

// Singular values (ascending) of the n x (n+1) upper bidiagonal with diagonal
// a[n] and super-diagonal b[n] by divide and conquer.
void dc_plan(ArenaPlan& p, int n);
int dc_run(int n, const double* a, const double* b, double* sigma, Arena& ws,
           hipStream_t st);

// Singular vectors of the n x n upper bidiagonal (a[n], b[n-1]) for the
// ascending singular values sigma (refined in place to high relative
// accuracy).  Vectors are written sigma-minor: Vt[j*ldv + i] = v_i[j] and
// Ut[j*ldv + i] = u_i[j], so B v_i = sigma_i u_i.
void twisted_plan(ArenaPlan& p, int n);
int twisted_run(int n, const double* a, const double* b, double* sigma,
                double* Vt, double* Ut, int ldv, Arena& ws, hipStream_t st);

// U = Q_left [Yb; 0] (m x n) and V = Q_right Xb (n x n) from the reflectors
// stored in A (m x n, lda), where Yb/Xb are given sigma-minor as from
// twisted_run (ld ldb).  U gets leading dimension ldu, V ldv.
void backtrans_plan(ArenaPlan& p, int m, int n);
int backtrans_run(int m, int n, const double* A, int lda, const double* Ut,
                  const double* Vt, int ldb, double* U, int ldu, double* V,
                  int ldv, Arena& ws, hipStream_t st);
// The same for m = n in FP32 arithmetic: reflectors stored as float in A32,
// the bidiagonal vectors Ut / Vt (FP64) rounded to float; U32, V32 (ld n).
void backtrans_plan_f32(ArenaPlan& p, int n);
int backtrans_run_f32(int n, const float* A32, int lda, const double* Ut, const double* Vt,
                      int ldb, float* U32, float* V32, Arena& ws, hipStream_t st);

// Explicit Q = H_0 ... H_{n-1} (left, m x m) or P = G_0 ... G_{n-3} (right,
// n x n) from the reflectors stored in A (m >= n).
void form_q_plan(ArenaPlan& p, int m, int n, bool left);
int form_q_run(bool left, int m, int n, const double* A, int lda, double* Q, int ldq, Arena& ws,
               hipStream_t st);

// Reverse-mode derivative of the thin SVD (m >= n): Abar from U, sigma, V and
// the cotangents (null = zero).  Device pointers.
void adjoint_plan(ArenaPlan& p, int m, int n);
int adjoint_run(int m, int n, const double* U, int ldu, const double* sigma,
                const double* V, int ldv, const double* Ub, int ldub, const double* sb,
                const double* Vb, int ldvb, double* Ab, int ldab, double guard,
                Arena& ws, hipStream_t st);

// D (c x r) = S^T for the r x c column-major matrix S.
int transpose_f64(int r, int c, const double* S, int lds, double* D, int ldd,
                  hipStream_t st);

// ---------------------------------------------------------------------------
// Library GEMMs (rocBLAS, MFMA), column-major, op 'N' or 'T'.  Stream-ordered.
// ---------------------------------------------------------------------------
int dgemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
          const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st);
int sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
          const float* B, int ldb, float beta, float* C, int ldc, hipStream_t st);
// BF16 operands (raw 16-bit storage), FP32 accumulation and C.
int gemm_bf16_f32(char ta, char tb, int M, int N, int K, float alpha, const void* A, int lda,
                  const void* B, int ldb, float beta, float* C, int ldc, hipStream_t st);
// The library's own FP64 GEMM on the matrix cores (gemm.hip), same
// arguments as dgemm.
int dgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
               const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st);
// The bidiagonalization's trailing update C (M x N) -= A B^T with inner
// dimension 64 (A: M x 64, ld lda; B: N x 64, ld ldb).  With cdot, also the
// column sums of the updated C against cdot (length M) over each 128-row
// tile I: part[I * ldp + j] for the columns j >= jfrom, in a fixed order.
int dgemm_trail64(int M, int N, const double* A, int lda, const double* B, int ldb, double* C,
                  int ldc, const double* cdot, double* part, int ldp, int jfrom, hipStream_t st);
constexpr int TRAIL_TILE = 128;  // rows per partial of dgemm_trail64
// B <- T^{-1} B for the M x M upper triangular T (non-unit diagonal).
int dtrsm_left_upper(int M, int N, const double* T, int ldt, double* B, int ldb,
                     hipStream_t st);
int strsm_left_upper(int M, int N, const float* T, int ldt, float* B, int ldb, hipStream_t st);
// inv (N x N, upper) = T^{-1} for the N x N upper triangular T (non-unit).
int dtrtri_upper(int N, const double* A, int lda, double* inv, int ldinv, hipStream_t st);
// Upper triangle of C (N x N) = A^T A for the K x N matrix A.
int dsyrk_upper_t(int N, int K, const double* A, int lda, double* C, int ldc, hipStream_t st);

// Deterministic wave64 sum (fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int B>
struct Log2;
template <>
struct Log2<2> {
  static constexpr int v = 1;
};
template <>
struct Log2<4> {
  static constexpr int v = 2;
};
template <>
struct Log2<8> {
  static constexpr int v = 3;
};
template <>
struct Log2<16> {
  static constexpr int v = 4;
};
template <>
struct Log2<32> {
  static constexpr int v = 5;
};

// Wave sum of each g[i], delivered to the lanes of group i (the 64/B lanes
// with lane >> (6 - log2 B) == i).  The first log2 B butterfly steps halve the
// set of values a lane carries (it keeps the half its lane bit selects and
// sends the other half), the remaining steps finish the sum inside the
// group: (B - 1) + (6 - log2 B) exchanges instead of 6 B.  Every lane of a
// group ends with the same value.
template <typename T, int B>
__device__ __forceinline__ T reduce_scatter(T (&g)[B]) {
  constexpr int LB = Log2<B>::v;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int L = 0; L < LB; ++L) {
    const int bit = 5 - L;
    const int half = B >> (L + 1);
    const bool upper = (lane >> bit) & 1;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      // the empty asm keeps both values in registers: without it the two
      // selects become one load from a selected address, which sends the
      // whole array to scratch memory
      T lo = g[j], hi = g[j + half];
      asm volatile("" : "+v"(lo), "+v"(hi));
      const T send = upper ? lo : hi;
      const T keep = upper ? hi : lo;
      g[j] = keep + __shfl_xor(send, 1 << bit, 64);
    }
  }
  T v = g[0];
#pragma unroll
  for (int bit = 5 - LB; bit >= 0; --bit) v += __shfl_xor(v, 1 << bit, 64);
  return v;
}

}  // namespace ddc
