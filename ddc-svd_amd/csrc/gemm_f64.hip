// gemm_f64.hip -- FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[M x N] = beta*C + alpha * sum_p op(A_p) op(B_p), column-major, up to two
// operand pairs (the blocked bidiagonal update A -= U Y^T + X V^T uses both,
// so C is read and written once).  Operands may be "structured": a mask on
// the stored indices zeroes the part of a reflector panel that the in-place
// storage uses for other data (see MASK_* in ddc_common.h).
//
// Tiling: 256 threads (4 wave64) per 128 x 128 C tile; each wave owns a
// 64 x 64 quadrant made of 4 x 4 MFMA 16 x 16 blocks (64 FP64 accumulators
// per lane).  K is staged through LDS in chunks of 16; the next chunk is
// fetched into registers (branch-free, predicated) while the MFMAs consume
// the current one.  Long-K products with few output tiles (W^T C in the
// back-transformation) are split over K: each split writes its partial
// tile, and a second kernel sums the splits in a fixed order, so results
// are bitwise reproducible.
#include <algorithm>

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
  int ksplit;      // number of K splits (grid.z)
  double* part;    // ksplit > 1: partial tiles [ksplit][M][N] (ld M)
};

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16;
constexpr int LDA_S = BK + 1;  // LDS row of the A tile (r-major: [BM][BK+1])
constexpr int LDB_S = BN + 1;  // LDS row of the B tile ([BK][BN+1])
constexpr int PER_T = BM * BK / 256;  // elements of each tile per thread (8)

__device__ __forceinline__ bool mask_ok(int mask, int moff, int i, int j) {
  if (mask == MASK_LOWER) return i >= j + moff;
  if (mask == MASK_UPPER1) return j >= i + moff;
  return true;
}

// op(X)(r, c), zero outside [0,R) x [0,Cc) or outside the mask; the load is
// always issued (at a clamped address) so the compiler can batch them.
__device__ __forceinline__ double fetch(const GemmOperand& X, int r, int c, int R, int Cc) {
  const bool in = (r < R) && (c < Cc);
  const int i = (X.op == 'N') ? r : c;
  const int j = (X.op == 'N') ? c : r;
  const bool ok = in && mask_ok(X.mask, X.moff, i, j);
  const int64_t off = ok ? ((int64_t)i + (int64_t)j * X.ld) : 0;
  double v = X.p[off];
  return ok ? v : 0.0;
}

// Thread -> element mapping for one K chunk; consecutive threads take
// consecutive stored addresses for either orientation.
__device__ __forceinline__ void a_coord(const GemmOperand& A, int e, int& r, int& c) {
  if (A.op == 'N') { r = e % BM; c = e / BM; }   // contiguous in r
  else             { c = e % BK; r = e / BK; }   // contiguous in k
}
__device__ __forceinline__ void b_coord(const GemmOperand& B, int e, int& k, int& c) {
  if (B.op == 'N') { k = e % BK; c = e / BK; }   // contiguous in k
  else             { c = e % BN; k = e / BN; }   // contiguous in n
}

__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g) {
  __shared__ double As[BM * LDA_S];
  __shared__ double Bs[BK * LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int row0 = blockIdx.x * BM;
  const int col0 = blockIdx.y * BN;
  const int wr = (wave & 1) * 64;   // wave quadrant row offset in the tile
  const int wc = (wave >> 1) * 64;  // wave quadrant column offset
  const int li = lane & 15;
  const int lk = lane >> 4;

  double4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};

  // K chunks of all pairs form one sequence; this split takes [cbeg, cend).
  int nch[2] = {0, 0};
  int total = 0;
  for (int p = 0; p < g.npairs; ++p) {
    nch[p] = (g.pair[p].K + BK - 1) / BK;
    total += nch[p];
  }
  const int per = (total + g.ksplit - 1) / g.ksplit;
  const int cbeg = blockIdx.z * per;
  const int cend = min(total, cbeg + per);

  double ra[PER_T], rb[PER_T];
  auto load_chunk = [&](int ch) {
    const int p = (ch < nch[0]) ? 0 : 1;
    const int k0 = (ch - (p ? nch[0] : 0)) * BK;
    const GemmOperand& A = g.pair[p].A;
    const GemmOperand& B = g.pair[p].B;
    const int K = g.pair[p].K;
#pragma unroll
    for (int q = 0; q < PER_T; ++q) {
      int r, c;
      a_coord(A, tid + 256 * q, r, c);
      ra[q] = fetch(A, row0 + r, k0 + c, g.M, K);
    }
#pragma unroll
    for (int q = 0; q < PER_T; ++q) {
      int k, c;
      b_coord(B, tid + 256 * q, k, c);
      rb[q] = fetch(B, k0 + k, col0 + c, K, g.N);
    }
  };
  auto store_chunk = [&](int ch) {
    const int p = (ch < nch[0]) ? 0 : 1;
#pragma unroll
    for (int q = 0; q < PER_T; ++q) {
      int r, c;
      a_coord(g.pair[p].A, tid + 256 * q, r, c);
      As[r * LDA_S + c] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < PER_T; ++q) {
      int k, c;
      b_coord(g.pair[p].B, tid + 256 * q, k, c);
      Bs[k * LDB_S + c] = rb[q];
    }
  };

  if (cbeg < cend) load_chunk(cbeg);
  for (int ch = cbeg; ch < cend; ++ch) {
    __syncthreads();
    store_chunk(ch);
    __syncthreads();
    if (ch + 1 < cend) load_chunk(ch + 1);  // in flight during the MFMAs
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = As[(wr + a * 16 + li) * LDA_S + kk + lk];
#pragma unroll
      for (int b = 0; b < 4; ++b) bf[b] = Bs[(kk + lk) * LDB_S + wc + b * 16 + li];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  }

  // D layout of the 16 x 16 block: lane j + 16*g holds rows g, g+4, g+8,
  // g+12 of column j (item q is row g + 4q).
  const int cj = lane & 15;
  const int rg = lane >> 4;
  const bool split = g.ksplit > 1;
  double* P = split ? g.part + (int64_t)blockIdx.z * g.M * g.N : nullptr;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = col0 + wc + b * 16 + cj;
      if (col >= g.N) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = row0 + wr + a * 16 + rg + 4 * q;
        if (row >= g.M) continue;
        if (split) {
          P[(int64_t)row + (int64_t)col * g.M] = acc[a][b][q];
        } else {
          double* cp = g.C + (int64_t)row + (int64_t)col * g.ldc;
          double v = g.alpha * acc[a][b][q];
          *cp = (g.beta == 0.0) ? v : g.beta * (*cp) + v;
        }
      }
    }
  }
}

// LDS-free variant: every wave owns a (16*TM) x (16*TN) tile and loads its
// MFMA fragments straight from global memory (L2-resident panels), two K
// steps ahead; the accumulators start from beta*C so the C tile is read
// while the first fragments are in flight.  No barriers.
template <int TM, int TN>
__global__ __launch_bounds__(256) void gemm_f64_wave_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15;
  const int lk = lane >> 4;
  // 2 x 2 waves per group
  const int row0 = blockIdx.x * (32 * TM) + (wave & 1) * (16 * TM);
  const int col0 = blockIdx.y * (32 * TN) + (wave >> 1) * (16 * TN);
  const bool split = g.ksplit > 1;

  double4_t acc[TM][TN];
  const int cj = lane & 15;
  const int rg = lane >> 4;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
      if (!split && g.beta != 0.0) {
        const int col = col0 + b * 16 + cj;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = row0 + a * 16 + rg + 4 * q;
          const bool ok = row < g.M && col < g.N;
          const int64_t off = ok ? (int64_t)row + (int64_t)col * g.ldc : 0;
          double v = g.C[off];
          acc[a][b][q] = ok ? g.beta * v : 0.0;
        }
      }
    }

  // K steps (4 wide) of all pairs form one sequence; this split takes its share
  int nst[2] = {0, 0};
  int total = 0;
  for (int p = 0; p < g.npairs; ++p) {
    nst[p] = (g.pair[p].K + 3) / 4;
    total += nst[p];
  }
  const int per = (total + g.ksplit - 1) / g.ksplit;
  const int sbeg = blockIdx.z * per;
  const int send = min(total, sbeg + per);
  const double alpha = split ? 1.0 : g.alpha;

  auto load = [&](int st, double* af, double* bf) {
    const int p = (st < nst[0]) ? 0 : 1;
    const int k = (st - (p ? nst[0] : 0)) * 4 + lk;
    const GemmOperand& A = g.pair[p].A;
    const GemmOperand& B = g.pair[p].B;
    const int K = g.pair[p].K;
#pragma unroll
    for (int a = 0; a < TM; ++a) af[a] = alpha * fetch(A, row0 + a * 16 + li, k, g.M, K);
#pragma unroll
    for (int b = 0; b < TN; ++b) bf[b] = fetch(B, k, col0 + b * 16 + li, K, g.N);
  };

  double af0[TM], bf0[TN], af1[TM], bf1[TN];
  if (sbeg < send) load(sbeg, af0, bf0);
  if (sbeg + 1 < send) load(sbeg + 1, af1, bf1);
  for (int st = sbeg; st < send; st += 2) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af0[a], bf0[b], acc[a][b], 0, 0, 0);
    if (st + 2 < send) load(st + 2, af0, bf0);
    if (st + 1 < send) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af1[a], bf1[b], acc[a][b], 0, 0, 0);
      if (st + 3 < send) load(st + 3, af1, bf1);
    }
  }

  double* P = split ? g.part + (int64_t)blockIdx.z * g.M * g.N : nullptr;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int col = col0 + b * 16 + cj;
      if (col >= g.N) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = row0 + a * 16 + rg + 4 * q;
        if (row >= g.M) continue;
        if (split) P[(int64_t)row + (int64_t)col * g.M] = acc[a][b][q];
        else g.C[(int64_t)row + (int64_t)col * g.ldc] = acc[a][b][q];
      }
    }
}

// C = beta C + alpha * sum_s part[s], splits summed in order.
__global__ void gemm_splitk_reduce(int M, int N, int ksplit, const double* part,
                                   double alpha, double beta, double* C, int ldc) {
  const int64_t total = (int64_t)M * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < ksplit; ++z) s += part[(int64_t)z * total + e];
    const int64_t col = e / M, row = e - col * M;
    double* cp = C + row + col * ldc;
    double v = alpha * s;
    *cp = (beta == 0.0) ? v : beta * (*cp) + v;
  }
}

int g_gemm_variant = 2;  // 0: LDS-tiled, 1: 64x64 wave tiles, 2: 32x64 wave tiles

static int gemm_launch(int M, int N, double alpha, const GemmPairDev* pairs,
                       int npairs, double beta, double* C, int ldc, int ksplit,
                       double* part, hipStream_t st) {
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
  g.ksplit = (ksplit > 1 && part) ? ksplit : 1;
  g.part = part;
  if (g_gemm_variant == 1) {
    dim3 grid((unsigned)ceil_div(M, 128), (unsigned)ceil_div(N, 128), (unsigned)g.ksplit);
    hipLaunchKernelGGL((gemm_f64_wave_kernel<4, 4>), grid, dim3(256), 0, st, g);
  } else if (g_gemm_variant == 2) {
    dim3 grid((unsigned)ceil_div(M, 64), (unsigned)ceil_div(N, 128), (unsigned)g.ksplit);
    hipLaunchKernelGGL((gemm_f64_wave_kernel<2, 4>), grid, dim3(256), 0, st, g);
  } else {
    dim3 grid((unsigned)ceil_div(M, BM), (unsigned)ceil_div(N, BN), (unsigned)g.ksplit);
    hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, st, g);
  }
  DDC_LAUNCHED();
  if (g.ksplit > 1) {
    int64_t total = (int64_t)M * N;
    unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(total, 256), 4096);
    hipLaunchKernelGGL(gemm_splitk_reduce, dim3(blocks), dim3(256), 0, st, M, N, g.ksplit,
                       part, alpha, beta, C, ldc);
    DDC_LAUNCHED();
  }
  return DDC_OK;
}

int gemm_split_count(int M, int N, int K, size_t part_cap) {
  // aim for >= 1024 work-groups, keep >= 512 of K per split
  int64_t tiles = ceil_div(M, BM) * ceil_div(N, BN);
  int64_t s = ceil_div(1024, tiles);
  s = std::min<int64_t>(s, std::max<int64_t>(1, K / 512));
  s = std::min<int64_t>(s, 64);
  while (s > 1 && (size_t)s * M * N > part_cap) --s;
  return (int)std::max<int64_t>(1, s);
}

int gemm_f64(int M, int N, int K, double alpha, GemmOperand A, GemmOperand B,
             double beta, double* C, int ldc, hipStream_t st) {
  GemmPairDev p{A, B, K};
  return gemm_launch(M, N, alpha, &p, 1, beta, C, ldc, 1, nullptr, st);
}

int gemm_f64_split(int M, int N, int K, double alpha, GemmOperand A, GemmOperand B,
                   double beta, double* C, int ldc, double* part, size_t part_cap,
                   hipStream_t st) {
  GemmPairDev p{A, B, K};
  int s = part ? gemm_split_count(M, N, K, part_cap) : 1;
  return gemm_launch(M, N, alpha, &p, 1, beta, C, ldc, s, part, st);
}

int gemm_f64_2(int M, int N, double alpha, GemmOperand A1, GemmOperand B1,
               int K1, GemmOperand A2, GemmOperand B2, int K2, double beta,
               double* C, int ldc, hipStream_t st) {
  GemmPairDev p[2] = {{A1, B1, K1}, {A2, B2, K2}};
  return gemm_launch(M, N, alpha, p, 2, beta, C, ldc, 1, nullptr, st);
}

}  // namespace ddc
