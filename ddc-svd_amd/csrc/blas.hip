// blas.hip -- plain dense GEMMs and the triangular solve through rocBLAS,
// whose FP64/FP32/BF16 kernels run on the MFMA units.  Used wherever a
// contraction is an ordinary dense product (after packing structured
// operands densely): the blocked trailing update, the compact-WY
// back-transformation, the adjoint and the refinement step.
#include <rocblas/rocblas.h>

#include <mutex>

#include "ddc_common.h"

namespace ddc {

namespace {
std::mutex g_blas_mu;
constexpr int kMaxDev = 64;
// One handle per (device, stream) for up to kSlots streams: a handle's
// workspace (split-K GEMMs, the triangular solve) must not be shared by
// calls that may run concurrently on different streams.  Further streams
// share slot 0, re-bound per call.
constexpr int kSlots = 4;
rocblas_handle g_handle[kMaxDev][kSlots] = {};
hipStream_t g_hstream[kMaxDev][kSlots] = {};

rocblas_operation op_of(char t) {
  return (t == 'T' || t == 't') ? rocblas_operation_transpose : rocblas_operation_none;
}

int blas_fail(const char* what, rocblas_status s) {
  char msg[160];
  snprintf(msg, sizeof(msg), "%s failed: %s", what, rocblas_status_to_string(s));
  report_error(msg);
  return DDC_E_HIP;
}
}  // namespace

// Handle of the current device bound to stream st; the caller holds g_blas_mu.
static int blas_handle(hipStream_t st, rocblas_handle* h) {
  int dev = 0;
  DDC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return DDC_E_ARG;
  int slot = -1;
  for (int i = 0; i < kSlots && slot < 0; ++i)
    if (g_handle[dev][i] && g_hstream[dev][i] == st) slot = i;
  for (int i = 0; i < kSlots && slot < 0; ++i)
    if (!g_handle[dev][i]) {
      rocblas_status s = rocblas_create_handle(&g_handle[dev][i]);
      if (s != rocblas_status_success) {
        g_handle[dev][i] = nullptr;
        return blas_fail("rocblas_create_handle", s);
      }
      slot = i;
    }
  if (slot < 0) slot = 0;
  if (g_hstream[dev][slot] != st || slot == 0) {
    rocblas_status s = rocblas_set_stream(g_handle[dev][slot], st);
    if (s != rocblas_status_success) return blas_fail("rocblas_set_stream", s);
    g_hstream[dev][slot] = st;
  }
  *h = g_handle[dev][slot];
  return DDC_OK;
}

int dgemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
          const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  rocblas_status s = rocblas_dgemm(h, op_of(ta), op_of(tb), M, N, K, &alpha, A, lda, B, ldb,
                                   &beta, C, ldc);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_dgemm", s);
}

int sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
          const float* B, int ldb, float beta, float* C, int ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  rocblas_status s = rocblas_sgemm(h, op_of(ta), op_of(tb), M, N, K, &alpha, A, lda, B, ldb,
                                   &beta, C, ldc);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_sgemm", s);
}

int gemm_bf16_f32(char ta, char tb, int M, int N, int K, float alpha, const void* A, int lda,
                  const void* B, int ldb, float beta, float* C, int ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  rocblas_status s = rocblas_gemm_ex(h, op_of(ta), op_of(tb), M, N, K, &alpha, A,
                                     rocblas_datatype_bf16_r, lda, B, rocblas_datatype_bf16_r,
                                     ldb, &beta, C, rocblas_datatype_f32_r, ldc, C,
                                     rocblas_datatype_f32_r, ldc, rocblas_datatype_f32_r,
                                     rocblas_gemm_algo_standard, 0, 0);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_gemm_ex(bf16)", s);
}

int dtrsm_left_upper(int M, int N, const double* T, int ldt, double* B, int ldb,
                     hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  const double one = 1.0;
  rocblas_status s = rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper,
                                   rocblas_operation_none, rocblas_diagonal_non_unit, M, N, &one,
                                   T, ldt, B, ldb);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_dtrsm", s);
}

int dtrtri_upper(int N, const double* A, int lda, double* inv, int ldinv, hipStream_t st) {
  if (N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  rocblas_status s = rocblas_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, N, A, lda,
                                    inv, ldinv);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_dtrtri", s);
}

int dsyrk_upper_t(int N, int K, const double* A, int lda, double* C, int ldc, hipStream_t st) {
  if (N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  const double one = 1.0, zero = 0.0;
  rocblas_status s = rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, N, K, &one,
                                   A, lda, &zero, C, ldc);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_dsyrk", s);
}

int strsm_left_upper(int M, int N, const float* T, int ldt, float* B, int ldb, hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  std::lock_guard<std::mutex> lk(g_blas_mu);
  rocblas_handle h;
  DDC_TRY(blas_handle(st, &h));
  const float one = 1.0f;
  rocblas_status s = rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper,
                                   rocblas_operation_none, rocblas_diagonal_non_unit, M, N, &one,
                                   T, ldt, B, ldb);
  return s == rocblas_status_success ? DDC_OK : blas_fail("rocblas_strsm", s);
}

}  // namespace ddc
