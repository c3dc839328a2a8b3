// gemm_f64.hip -- FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[M x N] = beta*C + alpha * sum_p op(A_p) op(B_p), column-major, up to two
// operand pairs (the blocked bidiagonal update A -= U Y^T + X V^T uses both,
// so C is read and written once).  Operands may be "structured": a mask on
// the stored indices zeroes the part of a reflector panel that the in-place
// storage uses for other data (see MASK_* in ddc_common.h).
//
// Tiling: 256 threads (4 wave64) per 64 x 64 C tile, each wave a 32 x 32
// quadrant made of 2 x 2 MFMA 16 x 16 blocks; K is staged through LDS in
// chunks of 16 with coalesced loads for either operand orientation.
#include "ddc_common.h"

namespace ddc {

struct GemmPairDev {
  GemmOperand A, B;
  int K;
};

struct GemmArgs {
  int M, N;
  double alpha, beta;
  double* C;
  int ldc;
  int npairs;
  GemmPairDev pair[2];
};

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int LDA_S = BK + 1;  // padded LDS row of the A tile (row-major r x k)
constexpr int LDB_S = BN + 1;  // padded LDS row of the B tile (k x n)

__device__ __forceinline__ bool mask_ok(int mask, int moff, int i, int j) {
  if (mask == MASK_LOWER) return i >= j + moff;
  if (mask == MASK_UPPER1) return j >= i + moff;
  return true;
}

// op(X)(r, c) from stored X; rows/cols outside [0,R) x [0,Cc) read as 0.
__device__ __forceinline__ double load_op(const GemmOperand& X, int r, int c,
                                          int R, int Cc) {
  if (r >= R || c >= Cc) return 0.0;
  int i = (X.op == 'N') ? r : c;
  int j = (X.op == 'N') ? c : r;
  if (!mask_ok(X.mask, X.moff, i, j)) return 0.0;
  return X.p[(int64_t)i + (int64_t)j * X.ld];
}

__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g) {
  __shared__ double As[BM * LDA_S];
  __shared__ double Bs[BK * LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int row0 = blockIdx.x * BM;
  const int col0 = blockIdx.y * BN;
  const int wr = (wave & 1) * 32;   // wave quadrant row offset in the tile
  const int wc = (wave >> 1) * 32;  // wave quadrant column offset

  double4_t acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};

  for (int p = 0; p < g.npairs; ++p) {
    const GemmOperand A = g.pair[p].A;
    const GemmOperand B = g.pair[p].B;
    const int K = g.pair[p].K;
    for (int k0 = 0; k0 < K; k0 += BK) {
      // A tile: BM x BK of op(A), starting at (row0, k0).
      if (A.op == 'N') {
        // stored column-major in r: consecutive threads take consecutive rows
        for (int e = tid; e < BM * BK; e += 256) {
          int r = e % BM, c = e / BM;
          As[r * LDA_S + c] = load_op(A, row0 + r, k0 + c, g.M, K);
        }
      } else {
        for (int e = tid; e < BM * BK; e += 256) {
          int c = e % BK, r = e / BK;
          As[r * LDA_S + c] = load_op(A, row0 + r, k0 + c, g.M, K);
        }
      }
      // B tile: BK x BN of op(B), starting at (k0, col0).
      if (B.op == 'N') {
        for (int e = tid; e < BK * BN; e += 256) {
          int k = e % BK, c = e / BK;
          Bs[k * LDB_S + c] = load_op(B, k0 + k, col0 + c, K, g.N);
        }
      } else {
        for (int e = tid; e < BK * BN; e += 256) {
          int c = e % BN, k = e / BN;
          Bs[k * LDB_S + c] = load_op(B, k0 + k, col0 + c, K, g.N);
        }
      }
      __syncthreads();
      const int li = lane & 15;
      const int lk = lane >> 4;
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        double af[2], bf[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) af[a] = As[(wr + a * 16 + li) * LDA_S + kk + lk];
#pragma unroll
        for (int b = 0; b < 2; ++b) bf[b] = Bs[(kk + lk) * LDB_S + wc + b * 16 + li];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
      }
      __syncthreads();
    }
  }

  // D layout of the 16 x 16 block: lane j + 16*g holds rows g, g+4, g+8,
  // g+12 of column j (item q is row g + 4q).
  const int cj = lane & 15;
  const int rg = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      int col = col0 + wc + b * 16 + cj;
      if (col >= g.N) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int row = row0 + wr + a * 16 + rg + 4 * q;
        if (row >= g.M) continue;
        double* cp = g.C + (int64_t)row + (int64_t)col * g.ldc;
        double v = g.alpha * acc[a][b][q];
        *cp = (g.beta == 0.0) ? v : g.beta * (*cp) + v;
      }
    }
  }
}

static int gemm_launch(int M, int N, double alpha, const GemmPairDev* pairs,
                       int npairs, double beta, double* C, int ldc,
                       hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.alpha = alpha;
  g.beta = beta;
  g.C = C;
  g.ldc = ldc;
  g.npairs = npairs;
  for (int p = 0; p < npairs; ++p) g.pair[p] = pairs[p];
  dim3 grid((unsigned)ceil_div(M, BM), (unsigned)ceil_div(N, BN));
  hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, st, g);
  DDC_LAUNCHED();
  return DDC_OK;
}

int gemm_f64(int M, int N, int K, double alpha, GemmOperand A, GemmOperand B,
             double beta, double* C, int ldc, hipStream_t st) {
  GemmPairDev p{A, B, K};
  return gemm_launch(M, N, alpha, &p, 1, beta, C, ldc, st);
}

int gemm_f64_2(int M, int N, double alpha, GemmOperand A1, GemmOperand B1,
               int K1, GemmOperand A2, GemmOperand B2, int K2, double beta,
               double* C, int ldc, hipStream_t st) {
  GemmPairDev p[2] = {{A1, B1, K1}, {A2, B2, K2}};
  return gemm_launch(M, N, alpha, p, 2, beta, C, ldc, st);
}

}  // namespace ddc
