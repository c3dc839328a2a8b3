// backtrans.hip -- U = Q_left [Yb; 0], V = Q_right Xb on the FP64 matrix cores.
//
// Replaces multU / multV (bidiag_par.c:990-1095) and the per-vector OpenMP
// loop of svd_gpu.c:117-121, which apply the reflectors one vector at a
// time (BLAS1).  Here the reflectors stored in A are grouped into blocks of
// NBB and applied in compact WY form,
//     H_k0 ... H_{k0+NBB-1} = I - W T W^T,   T upper triangular,
// for the unit-norm convention H = I - 2 w w^T (tau = 2).  T is never formed:
// its inverse is T^{-1} = I/2 + strict_upper(W^T W) (the identity behind
// the recursive T-factor of LAPACK dlarft), so a block costs
//     G = W^T W,  Z = W^T C,  Z = T Z (triangular solve with T^{-1}),
//     C -= W Z
// with the block's reflectors first packed into a dense zero-padded W; the
// three products are library GEMMs (rocBLAS, FP64 MFMA).
// Left reflectors are the columns A[k:m, k]; right reflectors are the rows
// A[k, k+1:n] (leading dimension n for V -- the reference indexes the
// transposed copy with m, bidiag_par.c:1034-1036, wrong when m != n).
#include <algorithm>

#include "ddc_common.h"

namespace ddc {

constexpr int NBB = 512;  // reflectors per compact-WY block

// C(0:n, 0:n) = Bt^T where Bt[j*ldb + i] holds vector i, entry j;
// rows n..rows-1 of C are zeroed.  TC: element type of C.
template <typename TC>
__global__ void bt_unpack(int n, int rows, const double* Bt, int ldb, TC* C, int ldc) {
  __shared__ double tile[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: i (vector), by: j (entry)
  int tx = threadIdx.x, ty = threadIdx.y;          // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    int j = by + r, i = bx + tx;
    tile[r][tx] = (j < n && i < n) ? Bt[(int64_t)j * ldb + i] : 0.0;
  }
  __syncthreads();
  // write C(j, i): consecutive threads take consecutive j (rows)
  for (int r = ty; r < 32; r += 8) {
    int i = bx + r, j = by + tx;
    if (i < n && j < rows) C[(int64_t)j + (int64_t)i * ldc] = (TC)((j < n) ? tile[tx][r] : 0.0);
  }
}

// D (c x r) = S^T for the r x c matrix S.
__global__ void k_transpose(int r, int c, const double* S, int lds, double* D, int ldd) {
  __shared__ double tile[32][33];
  int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  int tx = threadIdx.x, ty = threadIdx.y;
  for (int q = ty; q < 32; q += 8) {
    int i = r0 + tx, j = c0 + q;
    if (i < r && j < c) tile[q][tx] = S[(int64_t)i + (int64_t)j * lds];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    int j = c0 + tx, i = r0 + q;
    if (i < r && j < c) D[(int64_t)j + (int64_t)i * ldd] = tile[tx][q];
  }
}

int transpose_f64(int r, int c, const double* S, int lds, double* D, int ldd, hipStream_t st) {
  if (r <= 0 || c <= 0) return DDC_OK;
  dim3 g((unsigned)ceil_div(r, 32), (unsigned)ceil_div(c, 32));
  hipLaunchKernelGGL(k_transpose, g, dim3(32, 8), 0, st, r, c, S, lds, D, ldd);
  DDC_LAUNCHED();
  return DDC_OK;
}

// Dense copy Wd (rows x nb, ld ldw) of a block of reflectors, zeros
// outside their support.  left: Wd(i, c) = A[k0+i, k0+c] for i >= c.
// right: Wd(i, c) = A[k0+c, k0+1+i] for i >= c (rows of A, read through an
// LDS tile so that both sides stay coalesced).
template <bool RIGHT, typename T>
__global__ void bt_pack(int rows, int nb, const T* S, int lda, T* Wd, int ldw) {
  __shared__ T tile[32][33];
  const int i0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  if (!RIGHT) {
    for (int q = ty; q < 32; q += 8) {
      const int i = i0 + tx, c = c0 + q;
      if (i < rows && c < nb) Wd[i + (int64_t)c * ldw] = i >= c ? S[i + (int64_t)c * lda] : T(0);
    }
  } else {
    for (int q = ty; q < 32; q += 8) {  // tile[c][i] = S(c, i), coalesced over c
      const int c = c0 + tx, i = i0 + q;
      tile[q][tx] = (c < nb && i < rows && i >= c) ? S[c + (int64_t)i * lda] : T(0);
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
      const int i = i0 + tx, c = c0 + q;
      if (i < rows && c < nb) Wd[i + (int64_t)c * ldw] = tile[tx][q];
    }
  }
}

// T^{-1} = I/2 + strict_upper(G) (nb x nb, ld nb) for the block
// H_0 ... H_{nb-1} = I - W T W^T of unit-norm reflectors (tau = 2).
template <typename T>
__global__ void bt_tinv(int nb, T* G, T* Ti) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nb * nb) return;
  const int i = e % nb, j = e / nb;
  Ti[e] = i < j ? G[e] : (i == j ? T(0.5) : T(0));
  // G's strict lower part to zero: G then receives T (upper triangle only)
  // and is used as a full matrix
  if (i > j) G[e] = T(0);
}

// Library GEMM / triangular solve in the precision of the block.
static int gemm_t(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st) {
  return dgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
}
static int gemm_t(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
                  const float* B, int ldb, float beta, float* C, int ldc, hipStream_t st) {
  return sgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
}
static int trsm_t(int M, int N, const double* T, int ldt, double* B, int ldb, hipStream_t st) {
  return dtrsm_left_upper(M, N, T, ldt, B, ldb, st);
}
static int trsm_t(int M, int N, const float* T, int ldt, float* B, int ldb, hipStream_t st) {
  return strsm_left_upper(M, N, T, ldt, B, ldb, st);
}

// The left and right applications are independent: the right one runs on a
// second stream of the device, forked after the unpacking and joined before
// returning, so that the two sequences of block products (each ending in
// small, latency-bound GEMMs and triangular solves) overlap.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kMaxSideDev = 64;
static SideStream g_side[kMaxSideDev];

static int side_stream(SideStream** out) {
  int dev = 0;
  DDC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxSideDev) return DDC_E_ARG;
  SideStream& sd = g_side[dev];
  if (!sd.s) DDC_HIP(hipStreamCreateWithFlags(&sd.s, hipStreamNonBlocking));
  if (!sd.fork) DDC_HIP(hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming));
  if (!sd.join) DDC_HIP(hipEventCreateWithFlags(&sd.join, hipEventDisableTiming));
  *out = &sd;
  return DDC_OK;
}

void backtrans_plan(ArenaPlan& p, int m, int n) {
  p.add<double>((size_t)NBB * NBB * 2);  // G (then T), T^{-1}
  p.add<double>((size_t)NBB * n);        // Z
  p.add<double>((size_t)m * NBB);        // packed reflectors
  p.add<double>((size_t)NBB * n);        // T Z
  // the right-hand application's own set
  p.add<double>((size_t)NBB * NBB * 2);
  p.add<double>((size_t)NBB * n);
  p.add<double>((size_t)n * NBB);
  p.add<double>((size_t)NBB * n);
}

// Apply H_k0 ... H_k1-1 (blocks, last block first) to the rows of C:
// Z = W^T C, Z = T Z, C -= W Z.  With Z2 (FP64): T is formed once per block
// from T^{-1} (triangular inverse of the nb x nb matrix) and applied as a
// GEMM into Z2 (nb x ncol); else by a triangular solve with T^{-1} in place
// (the solve's blocked small products took about a third of the
// back-transformation at 8192^2).
// left: w_k = A[k:m, k];  right: w_k = A[k, k+1:n] acting on rows k+1..n-1.
template <typename T>
static int apply_blocks(bool left, int m, int n, const T* A, int lda, int nref, T* C, int ldc,
                        int ncol, T* G, T* Ti, T* Z, T* Wd, hipStream_t st, T* Z2 = nullptr) {
  const int nblk = (int)ceil_div(nref, NBB);
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * NBB;
    const int nb = std::min(NBB, nref - k0);
    const int rows = left ? m - k0 : n - k0 - 1;
    if (rows <= 0) continue;
    T* Csub = left ? C + k0 : C + k0 + 1;
    dim3 gp((unsigned)ceil_div(rows, 32), (unsigned)ceil_div(nb, 32));
    if (left)
      hipLaunchKernelGGL((bt_pack<false, T>), gp, dim3(32, 8), 0, st, rows, nb,
                         A + k0 + (int64_t)k0 * lda, lda, Wd, rows);
    else
      hipLaunchKernelGGL((bt_pack<true, T>), gp, dim3(32, 8), 0, st, rows, nb,
                         A + k0 + (int64_t)(k0 + 1) * lda, lda, Wd, rows);
    DDC_LAUNCHED();
    DDC_TRY(gemm_t('T', 'N', nb, nb, rows, T(1), Wd, rows, Wd, rows, T(0), G, nb, st));
    hipLaunchKernelGGL(bt_tinv<T>, dim3((unsigned)ceil_div(nb * nb, 256)), dim3(256), 0, st, nb,
                       G, Ti);
    DDC_LAUNCHED();
    DDC_TRY(gemm_t('T', 'N', nb, ncol, rows, T(1), Wd, rows, Csub, ldc, T(0), Z, nb, st));
    T* Zt = Z;
    if constexpr (sizeof(T) == 8) {
      if (Z2) {
        // T = (T^{-1})^{-1} into G (no longer needed), then Z2 = T Z
        DDC_TRY(dtrtri_upper(nb, Ti, nb, G, nb, st));
        DDC_TRY(gemm_t('N', 'N', nb, ncol, nb, T(1), G, nb, Z, nb, T(0), Z2, nb, st));
        Zt = Z2;
      }
    }
    if (Zt == Z) DDC_TRY(trsm_t(nb, ncol, Ti, nb, Z, nb, st));
    DDC_TRY(gemm_t('N', 'N', rows, ncol, nb, T(-1), Wd, rows, Zt, nb, T(1), Csub, ldc, st));
  }
  return DDC_OK;
}

int backtrans_run(int m, int n, const double* A, int lda, const double* Ut,
                  const double* Vt, int ldb, double* U, int ldu, double* V,
                  int ldv, Arena& ws, hipStream_t st) {
  if (m < n || n < 1) return DDC_E_ARG;
  double* G = ws.take<double>((size_t)NBB * NBB * 2);
  double* Ti = G ? G + NBB * NBB : nullptr;
  double* Z = ws.take<double>((size_t)NBB * n);
  double* Wd = ws.take<double>((size_t)m * NBB);
  double* TZ = ws.take<double>((size_t)NBB * n);
  double* G2 = ws.take<double>((size_t)NBB * NBB * 2);
  double* Ti2 = G2 ? G2 + NBB * NBB : nullptr;
  double* Z2 = ws.take<double>((size_t)NBB * n);
  double* Wd2 = ws.take<double>((size_t)n * NBB);
  double* TZ2 = ws.take<double>((size_t)NBB * n);
  if (!TZ2) return DDC_E_NOMEM;
  SideStream* sd = nullptr;
  DDC_TRY(side_stream(&sd));
  dim3 tb(32, 8);
  dim3 gu((unsigned)ceil_div(n, 32), (unsigned)ceil_div(m, 32));
  hipLaunchKernelGGL(bt_unpack<double>, gu, tb, 0, st, n, m, Ut, ldb, U, ldu);
  DDC_LAUNCHED();
  dim3 gv((unsigned)ceil_div(n, 32), (unsigned)ceil_div(n, 32));
  hipLaunchKernelGGL(bt_unpack<double>, gv, tb, 0, st, n, n, Vt, ldb, V, ldv);
  DDC_LAUNCHED();
  const bool right = n > 2;
  if (right) {
    DDC_HIP(hipEventRecord(sd->fork, st));
    DDC_HIP(hipStreamWaitEvent(sd->s, sd->fork, 0));
  }
  // left: reflectors 0..n-1 (the last one is the 1 x 1 sign flip when m == n)
  const int rcl = apply_blocks(true, m, n, A, lda, n, U, ldu, n, G, Ti, Z, Wd, st, TZ);
  // right: reflectors 0..n-3 (row n-2 holds no reflection); joined even on
  // failure, so that nothing of this call still runs when it returns
  int rcr = DDC_OK;
  if (right) {
    rcr = apply_blocks(false, m, n, A, lda, n - 2, V, ldv, n, G2, Ti2, Z2, Wd2, sd->s, TZ2);
    DDC_HIP(hipEventRecord(sd->join, sd->s));
    DDC_HIP(hipStreamWaitEvent(st, sd->join, 0));
  }
  DDC_TRY(rcl);
  DDC_TRY(rcr);
  return DDC_OK;
}

void backtrans_plan_f32(ArenaPlan& p, int n) {
  for (int side = 0; side < 2; ++side) {
    p.add<float>((size_t)NBB * NBB * 2);
    p.add<float>((size_t)NBB * n);
    p.add<float>((size_t)n * NBB);
  }
}

// FP32 arithmetic throughout (the mixed-precision form refines the result).
int backtrans_run_f32(int n, const float* A32, int lda, const double* Ut, const double* Vt,
                      int ldb, float* U32, float* V32, Arena& ws, hipStream_t st) {
  if (n < 1) return DDC_E_ARG;
  float* G = ws.take<float>((size_t)NBB * NBB * 2);
  float* Ti = G ? G + NBB * NBB : nullptr;
  float* Z = ws.take<float>((size_t)NBB * n);
  float* Wd = ws.take<float>((size_t)n * NBB);
  float* G2 = ws.take<float>((size_t)NBB * NBB * 2);
  float* Ti2 = G2 ? G2 + NBB * NBB : nullptr;
  float* Z2 = ws.take<float>((size_t)NBB * n);
  float* Wd2 = ws.take<float>((size_t)n * NBB);
  if (!Wd2) return DDC_E_NOMEM;
  SideStream* sd = nullptr;
  DDC_TRY(side_stream(&sd));
  dim3 tb(32, 8);
  dim3 g((unsigned)ceil_div(n, 32), (unsigned)ceil_div(n, 32));
  hipLaunchKernelGGL(bt_unpack<float>, g, tb, 0, st, n, n, Ut, ldb, U32, n);
  hipLaunchKernelGGL(bt_unpack<float>, g, tb, 0, st, n, n, Vt, ldb, V32, n);
  DDC_LAUNCHED();
  const bool right = n > 2;
  if (right) {
    DDC_HIP(hipEventRecord(sd->fork, st));
    DDC_HIP(hipStreamWaitEvent(sd->s, sd->fork, 0));
  }
  const int rcl = apply_blocks<float>(true, n, n, A32, lda, n, U32, n, n, G, Ti, Z, Wd, st);
  int rcr = DDC_OK;
  if (right) {
    rcr = apply_blocks<float>(false, n, n, A32, lda, n - 2, V32, n, n, G2, Ti2, Z2, Wd2, sd->s);
    DDC_HIP(hipEventRecord(sd->join, sd->s));
    DDC_HIP(hipStreamWaitEvent(st, sd->join, 0));
  }
  DDC_TRY(rcl);
  DDC_TRY(rcr);
  return DDC_OK;
}

// C (rows x cols, ld ldc) = the leading part of the identity.
__global__ void bt_identity(int rows, int cols, double* C, int ldc) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / rows, i = e - j * rows;
    C[i + j * ldc] = i == j ? 1.0 : 0.0;
  }
}

void form_q_plan(ArenaPlan& p, int m, int n, bool left) {
  const int ncol = left ? m : n;
  p.add<double>((size_t)NBB * NBB * 2);
  p.add<double>((size_t)NBB * ncol);
  p.add<double>((size_t)(left ? m : n) * NBB);
}

// Q = H_0 ... H_{n-1} (left, m x m) or G_0 ... G_{n-3} (right, n x n) from the
// reflectors stored in A (m x n, m >= n), as form_u_par / form_v_par build
// them column by column (bidiag_par.c:877-988); here one compact-WY block at
// a time on the identity.
int form_q_run(bool left, int m, int n, const double* A, int lda, double* Q, int ldq, Arena& ws,
               hipStream_t st) {
  if (m < n || n < 1) return DDC_E_ARG;
  const int k = left ? m : n;
  double* G = ws.take<double>((size_t)NBB * NBB * 2);
  double* Ti = G ? G + NBB * NBB : nullptr;
  double* Z = ws.take<double>((size_t)NBB * k);
  double* Wd = ws.take<double>((size_t)k * NBB);
  if (!Wd) return DDC_E_NOMEM;
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div((int64_t)k * k, 256), 8192);
  hipLaunchKernelGGL(bt_identity, dim3(g), dim3(256), 0, st, k, k, Q, ldq);
  DDC_LAUNCHED();
  if (left) return apply_blocks(true, m, n, A, lda, n, Q, ldq, k, G, Ti, Z, Wd, st);
  if (n > 2) return apply_blocks(false, m, n, A, lda, n - 2, Q, ldq, k, G, Ti, Z, Wd, st);
  return DDC_OK;
}

}  // namespace ddc
