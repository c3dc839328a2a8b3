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
//     Z = W^T C  (MFMA GEMM),  Z = T Z by one upper-triangular solve per
//     column of Z,  C -= W Z  (MFMA GEMM).
// Left reflectors are the columns A[k:m, k]; right reflectors are the rows
// A[k, k+1:n] (leading dimension n for V -- the reference indexes the
// transposed copy with m, bidiag_par.c:1034-1036, wrong when m != n).
#include <algorithm>

#include "ddc_common.h"

namespace ddc {

constexpr int NBB = 64;

// C(0:n, 0:n) = Bt^T where Bt[j*ldb + i] holds vector i, entry j;
// rows n..rows-1 of C are zeroed.
__global__ void bt_unpack(int n, int rows, const double* Bt, int ldb, double* C, int ldc) {
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
    if (i < n && j < rows) C[(int64_t)j + (int64_t)i * ldc] = (j < n) ? tile[tx][r] : 0.0;
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

// Z <- T Z with T^{-1} = I/2 + strict_upper(G), G = W^T W (nb x nb, ld nb):
// back substitution z_i = 2 (z_i - sum_{j>i} G(i,j) z_j), one thread per
// column of Z (ld nb).
__global__ __launch_bounds__(64) void bt_trsm(int nb, int ncol, const double* G, double* Z) {
  __shared__ double Gs[NBB * NBB];
  for (int e = threadIdx.x; e < nb * nb; e += 64) Gs[e] = G[e];
  __syncthreads();
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= ncol) return;
  double z[NBB];
  double* zc = Z + (int64_t)c * nb;
#pragma unroll
  for (int i = 0; i < NBB; ++i) z[i] = (i < nb) ? zc[i] : 0.0;
#pragma unroll
  for (int i = NBB - 1; i >= 0; --i) {
    if (i >= nb) continue;
    double acc = z[i];
#pragma unroll
    for (int j = i + 1; j < NBB; ++j)
      if (j < nb) acc -= Gs[i + j * nb] * z[j];
    z[i] = 2.0 * acc;
  }
#pragma unroll
  for (int i = 0; i < NBB; ++i)
    if (i < nb) zc[i] = z[i];
}

constexpr int SPLITS = 16;

void backtrans_plan(ArenaPlan& p, int m, int n) {
  p.add<double>((size_t)NBB * NBB * 2);        // G, T
  p.add<double>((size_t)NBB * n * 2);          // Z, Z2
  p.add<double>((size_t)SPLITS * NBB * n);     // split-K partials
}

// Apply H_k0 ... H_k1-1 (blocks, last block first) to the rows of C.
// left: w_k = A[k:m, k];  right: w_k = A[k, k+1:n] acting on rows k+1..n-1.
static int apply_blocks(bool left, int m, int n, const double* A, int lda,
                        int nref, double* C, int ldc, int ncol, double* G,
                        double* T, double* Z, double* Z2, double* part,
                        size_t part_cap, hipStream_t st) {
  const int nblk = (int)ceil_div(nref, NBB);
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * NBB;
    const int nb = std::min(NBB, nref - k0);
    GemmOperand W, Wt;
    int rows;
    double* Csub;
    if (left) {
      // W = A[k0:m, k0:k0+nb], w_c has support rows >= c (local)
      rows = m - k0;
      W = GemmOperand{A + k0 + (int64_t)k0 * lda, lda, 'N', MASK_LOWER, 0};
      Csub = C + k0;
    } else {
      // R = A[k0:k0+nb, k0+1:n] (row storage), W = R^T; w_c has support
      // local columns >= c
      rows = n - k0 - 1;
      W = GemmOperand{A + k0 + (int64_t)(k0 + 1) * lda, lda, 'T', MASK_UPPER1, 0};
      Csub = C + k0 + 1;
    }
    if (rows <= 0) continue;
    Wt = W;
    Wt.op = (W.op == 'N') ? 'T' : 'N';
    // G = W^T W ; Z = W^T C ; Z = T Z ; C -= W Z
    DDC_TRY(gemm_f64_split(nb, nb, rows, 1.0, Wt, W, 0.0, G, nb, part, part_cap, st));
    GemmOperand Cop{Csub, ldc, 'N', MASK_NONE, 0};
    DDC_TRY(gemm_f64_split(nb, ncol, rows, 1.0, Wt, Cop, 0.0, Z, nb, part, part_cap, st));
    hipLaunchKernelGGL(bt_trsm, dim3((unsigned)ceil_div(ncol, 64)), dim3(64), 0, st, nb, ncol,
                       G, Z);
    DDC_LAUNCHED();
    GemmOperand Z2op{Z, nb, 'N', MASK_NONE, 0};
    DDC_TRY(gemm_f64(rows, ncol, nb, -1.0, W, Z2op, 1.0, Csub, ldc, st));
  }
  return DDC_OK;
}

int backtrans_run(int m, int n, const double* A, int lda, const double* Ut,
                  const double* Vt, int ldb, double* U, int ldu, double* V,
                  int ldv, Arena& ws, hipStream_t st) {
  if (m < n || n < 1) return DDC_E_ARG;
  double* G = ws.take<double>((size_t)NBB * NBB * 2);
  double* T = G + NBB * NBB;
  double* Z = ws.take<double>((size_t)NBB * n * 2);
  double* Z2 = Z + (size_t)NBB * n;
  const size_t part_cap = (size_t)SPLITS * NBB * n;
  double* part = ws.take<double>(part_cap);
  if (!part) return DDC_E_NOMEM;
  dim3 tb(32, 8);
  dim3 gu((unsigned)ceil_div(n, 32), (unsigned)ceil_div(m, 32));
  hipLaunchKernelGGL(bt_unpack, gu, tb, 0, st, n, m, Ut, ldb, U, ldu);
  DDC_LAUNCHED();
  dim3 gv((unsigned)ceil_div(n, 32), (unsigned)ceil_div(n, 32));
  hipLaunchKernelGGL(bt_unpack, gv, tb, 0, st, n, n, Vt, ldb, V, ldv);
  DDC_LAUNCHED();
  // left: reflectors 0..n-1 (the last one is the 1 x 1 sign flip when m == n)
  DDC_TRY(apply_blocks(true, m, n, A, lda, n, U, ldu, n, G, T, Z, Z2, part, part_cap, st));
  // right: reflectors 0..n-3 (row n-2 holds no reflection)
  if (n > 2)
    DDC_TRY(apply_blocks(false, m, n, A, lda, n - 2, V, ldv, n, G, T, Z, Z2, part, part_cap, st));
  return DDC_OK;
}

}  // namespace ddc
