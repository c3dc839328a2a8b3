// gemm.hip -- FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
//   C = alpha op(A) op(B) + beta C      (column-major, op = N or T)
//
// The dense contractions of the bidiagonalization (the blocked trailing
// update A0 -= [U X] [Y V]^T, inner dimension 2 NB: k_trail64 below) and
// of the back-transformation; no reference counterpart (the reference
// applies every rank-1 update separately, left_update_mat.cl:33-49,
// right_update_mat.cl:33-49).
//
// Tiling of the general kernel (k_dgemm).  A work-group of 256 threads (four waves, 2 x 2) owns a 128 x 128
// tile of C; each wave a 64 x 64 quarter, i.e. 4 x 4 MFMA blocks of 16 x 16
// whose accumulators (four doubles per lane each, 64 per lane) stay in
// registers for the whole K loop.  The accumulators start from beta C (its
// loads are in flight together with those of the first K slice), and op(A)
// is scaled by alpha as it is staged, so D = (alpha op(A)) op(B) + beta C
// needs no read of C after the products.  op(A) and op(B) advance through
// LDS in slices of KS = 16 (128 x 16 each, 32 KiB per stage, two stages), so
// each operand value loaded from global memory feeds two waves.
//
// Operand layouts of v_mfma_f64_16x16x4_f64 (wave64), lane l (measured on
// gfx950 with single-entry operands):
//   A block (16 x 4):  A[l % 16][l / 16]
//   B block (4 x 16):  B[l / 16][l % 16]
//   D block (16 x 16): D[l / 16 + 4 r][l % 16], r = 0..3 (the double4)
// The A rows are fed permuted, row p(x) = 4 (x % 4) + x / 4 in slot x, so
// that D row l / 16 + 4 r holds C row 4 (l / 16) + r: a lane's four results
// are four consecutive rows of one column of C, one 32-byte load and store
// per block in column-major C.
#include <algorithm>
#include <cstdlib>

#include "ddc_common.h"

namespace ddc {

namespace {

constexpr int GT = 128;  // C tile (rows and columns) per work-group
constexpr int KS = 16;   // K slice per LDS stage
typedef double d4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
  int M, N, K;
  double alpha, beta;
  const double* A;
  int lda;
  const double* B;
  int ldb;
  double* C;
  int ldc;
};

// A thread's share of one K slice [k0, k0 + KS) of op(X) for the 128 rows
// (A) or columns (B) starting at r0: fetch() issues its global loads into
// registers, put() writes them, times scale, to S[k][r] (r contiguous), so
// the loads of the next slice are in flight while the current one is
// multiplied.  TX: X is stored transposed (op(X)[r][k] = X[k + r ldx]),
// else op(X)[r][k] = X[r + k ldx].  Entries outside [0, R) x [0, K) are
// zero.
template <bool TX, int NT>
struct Slice {
  static constexpr int NV = 1024 / NT;  // double2 values per thread
  double2 v[NV];
  __device__ __forceinline__ void fetch(const double* X, int ldx, int R, int K, int r0, int k0) {
    const int tid = threadIdx.x;
    if (!TX) {
      // 128 rows x 16 k: thread = (row pair, run of NV k), double2 loads along r
      const int rp = (tid & 63) * 2, kq = (tid >> 6) * NV;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int k = k0 + kq + u, r = r0 + rp;
        v[u] = make_double2(0.0, 0.0);
        if (k < K) {
          const double* p = X + (int64_t)k * ldx + r;
          if (r + 1 < R && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
            v[u] = *reinterpret_cast<const double2*>(p);
          } else {
            if (r < R) v[u].x = p[0];
            if (r + 1 < R) v[u].y = p[1];
          }
        }
      }
    } else {
      // 16 k x 128 rows, k contiguous: thread = (row, run of 2 NV k), double2 loads
      const int r = tid & 127, kq = (tid >> 7) * 2 * NV, rr = r0 + r;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int k = k0 + kq + 2 * u;
        v[u] = make_double2(0.0, 0.0);
        if (rr < R) {
          const double* p = X + (int64_t)rr * ldx + k;
          if (k + 1 < K && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
            v[u] = *reinterpret_cast<const double2*>(p);
          } else {
            if (k < K) v[u].x = p[0];
            if (k + 1 < K) v[u].y = p[1];
          }
        }
      }
    }
  }
  __device__ __forceinline__ void put(double (*S)[GT], double scale) const {
    const int tid = threadIdx.x;
    if (!TX) {
      const int rp = (tid & 63) * 2, kq = (tid >> 6) * NV;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        S[kq + u][rp] = scale * v[u].x;
        S[kq + u][rp + 1] = scale * v[u].y;
      }
    } else {
      const int r = tid & 127, kq = (tid >> 7) * 2 * NV;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        S[kq + 2 * u][r] = scale * v[u].x;
        S[kq + 2 * u + 1][r] = scale * v[u].y;
      }
    }
  }
};

// NW waves: NW = 4 as 2 x 2 waves of 64 x 64, NW = 8 as 4 x 2 waves of
// 32 x 64 (half the accumulators per wave, twice the waves to hide latency)
template <bool TA, bool TB, int NW>
__global__ __launch_bounds__(64 * NW, 2) void k_dgemm(GemmArgs g) {
  constexpr int NT = 64 * NW;
  constexpr int WR = GT / (NW / 2);  // rows per wave
  constexpr int MA = WR / 16;        // 16-row blocks per wave
  __shared__ double As[2][KS][GT];
  __shared__ double Bs[2][KS][GT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
  const int wi = (wave % (NW / 2)) * WR, wj = (wave / (NW / 2)) * 64;
  const int lr = lane & 15, lq = lane >> 4;
  const int lp = 4 * (lr & 3) + (lr >> 2);  // permuted A row of this lane's slot
  // 32-byte accesses to C when every column start is 32-byte aligned
  const bool vec4 = ((reinterpret_cast<uintptr_t>(g.C) | ((uint64_t)g.ldc * 8)) & 31) == 0;

  d4 acc[MA][4];
#pragma unroll
  for (int a = 0; a < MA; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      const int i = i0 + wi + 16 * a + 4 * lq;
      const int j = j0 + wj + 16 * b + lr;
      if (g.beta == 0.0 || j >= g.N || i >= g.M) continue;
      const double* c = g.C + i + (int64_t)j * g.ldc;
      if (vec4 && i + 3 < g.M) {
        acc[a][b] = g.beta * *reinterpret_cast<const d4*>(c);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (i + r < g.M) acc[a][b][r] = g.beta * c[r];
      }
    }

  const int nk = (g.K + KS - 1) / KS;
  Slice<TA, NT> sa;
  Slice<TB, NT> sb;
  sa.fetch(g.A, g.lda, g.M, g.K, i0, 0);
  sb.fetch(g.B, g.ldb, g.N, g.K, j0, 0);
  sa.put(As[0], g.alpha);
  sb.put(Bs[0], 1.0);
  __syncthreads();
  for (int s = 0; s < nk; ++s) {
    const int cur = s & 1;
    const bool more = s + 1 < nk;
    if (more) {
      sa.fetch(g.A, g.lda, g.M, g.K, i0, (s + 1) * KS);
      sb.fetch(g.B, g.ldb, g.N, g.K, j0, (s + 1) * KS);
    }
#pragma unroll
    for (int kk = 0; kk < KS; kk += 4) {
      double av[MA], bv[4];
#pragma unroll
      for (int a = 0; a < MA; ++a) av[a] = As[cur][kk + lq][wi + 16 * a + lp];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[cur][kk + lq][wj + 16 * b + lr];
#pragma unroll
      for (int a = 0; a < MA; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    // the other buffer was last read before the previous barrier
    if (more) {
      sa.put(As[cur ^ 1], g.alpha);
      sb.put(Bs[cur ^ 1], 1.0);
    }
    __syncthreads();
  }

  // store: the lane holds rows 4 lq .. 4 lq + 3 of column lr of each block
#pragma unroll
  for (int a = 0; a < MA; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + wi + 16 * a + 4 * lq;
      const int j = j0 + wj + 16 * b + lr;
      if (j >= g.N || i >= g.M) continue;
      double* c = g.C + i + (int64_t)j * g.ldc;
      if (vec4 && i + 3 < g.M) {
        *reinterpret_cast<d4*>(c) = acc[a][b];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (i + r < g.M) c[r] = acc[a][b][r];
      }
    }
}

template <int NW>
static void launch_dgemm(bool TA, bool TB, const GemmArgs& g, hipStream_t st) {
  const dim3 grid((unsigned)ceil_div(g.M, GT), (unsigned)ceil_div(g.N, GT));
  const dim3 block(64 * NW);
  // op(B)[k][j]: B stored N (K x N, B[k + j ldb]) reads like a transposed
  // A, and stored T (N x K, B[j + k ldb]) like a plain A
  if (!TA && !TB) hipLaunchKernelGGL((k_dgemm<false, true, NW>), grid, block, 0, st, g);
  if (!TA && TB) hipLaunchKernelGGL((k_dgemm<false, false, NW>), grid, block, 0, st, g);
  if (TA && !TB) hipLaunchKernelGGL((k_dgemm<true, true, NW>), grid, block, 0, st, g);
  if (TA && TB) hipLaunchKernelGGL((k_dgemm<true, false, NW>), grid, block, 0, st, g);
}

// ---- the trailing update of the bidiagonalization, inner dimension 64 ------
//
// C -= A B^T with K = 64 moves 16 bytes of C per 128 flops: on gfx950 the
// FP64 matrix cores need about as long for a 128 x 128 tile (6.8 us at
// 307 GFLOP/s per CU) as HBM needs for its C traffic at the CU's share of
// 8 TB/s, so the two must overlap.  One work-group of eight waves per CU:
// both operand panels of the tile (128 x 64 each, 128 KiB) are staged in
// LDS at once (one barrier, no K loop), and the tile's C is loaded into
// registers of its own at the start and added only after the products,
// so the C loads are in flight during the whole MFMA phase.
//
// With cdot the epilogue also forms, for each column j of the tile, the
// sum over the tile's rows of the updated C[i][j] times cdot[i], in a fixed
// order (lanes, then waves): the partial products A0^T c of the next
// column step's left GEMV, taken while the tile is in registers.

struct TrailArgs {
  int M, N;
  double alpha, beta;
  const double* A;
  int lda;
  const double* B;
  int ldb;
  double* C;
  int ldc;
  const double* cdot;
  double* part;
  int ldp, jfrom;
};

// Rows r0 .. r0+127 of the M x 64 panel X (column-major, ld ldx) as
// S[k][r], times scale; rows past R are zero.  Thread = (row pair, k run
// with stride NT / 64), double2 loads along the rows.
template <int NT>
struct Panel64 {
  static constexpr int NV = 64 * GT / (2 * NT), KSTEP = NT / 64;
  double2 v[NV];
  __device__ __forceinline__ void fetch(const double* X, int ldx, int R, int r0, bool a16) {
    const int tid = threadIdx.x, rp = (tid & 63) * 2, kb = tid >> 6, r = r0 + rp;
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const double* p = X + (int64_t)(kb + KSTEP * u) * ldx + r;
      if (a16 && r + 1 < R) {
        v[u] = *reinterpret_cast<const double2*>(p);
      } else {
        v[u].x = r < R ? p[0] : 0.0;
        v[u].y = r + 1 < R ? p[1] : 0.0;
      }
    }
  }
  __device__ __forceinline__ void put(double (*S)[GT], double scale) const {
    const int tid = threadIdx.x, rp = (tid & 63) * 2, kb = tid >> 6;
#pragma unroll
    for (int u = 0; u < NV; ++u)
      *reinterpret_cast<double2*>(&S[kb + KSTEP * u][rp]) =
          make_double2(scale * v[u].x, scale * v[u].y);
  }
};

// A wave's C blocks of the tile (i0, j0): rows i0 + wi + 16 a + 4 lq + r of
// the columns j0 + wj + 16 b + lr (zero outside the matrix, or if beta = 0).
template <int NBB>
__device__ __forceinline__ void load_cblk(const TrailArgs& g, int i0, int j0, int wi, int wj,
                                          int lr, int lq, bool vec4, d4 (&cv)[2][NBB]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NBB; ++b) {
      cv[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      const int i = i0 + wi + 16 * a + 4 * lq;
      const int j = j0 + wj + 16 * b + lr;
      if (g.beta == 0.0 || j >= g.N || i >= g.M) continue;
      const double* c = g.C + i + (int64_t)j * g.ldc;
      if (vec4 && i + 3 < g.M) {
        cv[a][b] = *reinterpret_cast<const d4*>(c);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (i + r < g.M) cv[a][b][r] = c[r];
      }
    }
}

// Persistent: work-group w takes the tiles [w T / G, (w + 1) T / G) of the
// T tiles in row-major order (consecutive tiles share their A panel, which
// then stays in LDS), one work-group per CU.  Per tile: the accumulators
// start from beta C (loaded during the previous tile's products); the
// loads of the next tile's C and B panel are issued before this tile's
// products, and its stores after them, so that HBM is busy while the
// matrix cores are.  A new tile row (at most once or twice per work-group
// when it has fewer tiles than a row) loads its A panel after the products.
template <bool DOT>
__global__ __launch_bounds__(1024, 1) void k_trail64(TrailArgs g, int tn, int ntile) {
  // sixteen waves as 4 x 4 of 32 x 32 (four waves per SIMD to keep the
  // matrix cores fed; eight waves of 32 x 64: 0.314 against 0.271 ms for
  // an 8160^2 update)
  constexpr int KT = 64, MA = 2, NBB = 2, NT = 1024;
  __shared__ double As[KT][GT];
  __shared__ double Bs[KT][GT];
  __shared__ double red[4][GT];
  __shared__ double cdl[GT];  // cdot on the tile row
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 3) * 32, wj = (wave >> 2) * 32;
  const int lr = lane & 15, lq = lane >> 4;
  const int lp = 4 * (lr & 3) + (lr >> 2);
  const bool vec4 = ((reinterpret_cast<uintptr_t>(g.C) | ((uint64_t)g.ldc * 8)) & 31) == 0;
  const bool a16 = ((reinterpret_cast<uintptr_t>(g.A) | ((uint64_t)g.lda * 8)) & 15) == 0;
  const bool b16 = ((reinterpret_cast<uintptr_t>(g.B) | ((uint64_t)g.ldb * 8)) & 15) == 0;
  const int G = (int)gridDim.x, w = (int)blockIdx.x;
  const int t0 = (int)((int64_t)w * ntile / G), t1 = (int)((int64_t)(w + 1) * ntile / G);
  if (t0 >= t1) return;  // uniform over the work-group

  int I = t0 / tn, J = t0 % tn;
  d4 cn[MA][NBB];
  {
    Panel64<NT> pa;
    pa.fetch(g.A, g.lda, g.M, I * GT, a16);
    pa.put(As, g.alpha);
  }
  Panel64<NT> pb;
  pb.fetch(g.B, g.ldb, g.N, J * GT, b16);
  load_cblk<NBB>(g, I * GT, J * GT, wi, wj, lr, lq, vec4, cn);
  if (DOT && tid < GT) cdl[tid] = I * GT + tid < g.M ? g.cdot[I * GT + tid] : 0.0;
  pb.put(Bs, 1.0);
  __syncthreads();

  for (int n = t0; n < t1; ++n) {
    const int i0 = I * GT, j0 = J * GT;
    d4 acc[MA][NBB];
#pragma unroll
    for (int a = 0; a < MA; ++a)
#pragma unroll
      for (int b = 0; b < NBB; ++b) acc[a][b] = g.beta * cn[a][b];
    // the next tile's loads, in flight during this tile's products
    const bool more = n + 1 < t1;
    const int In = (n + 1) / tn, Jn = (n + 1) % tn;
    const bool newrow = more && In != I;
    if (more) {
      load_cblk<NBB>(g, In * GT, Jn * GT, wi, wj, lr, lq, vec4, cn);
      pb.fetch(g.B, g.ldb, g.N, Jn * GT, b16);
    }
// (two k steps at a time with the column sums: four spill registers)
    constexpr int UNR = DOT ? 2 : 4;
#pragma unroll UNR
    for (int kk = 0; kk < KT; kk += 4) {
      double av[MA], bv[NBB];
#pragma unroll
      for (int a = 0; a < MA; ++a) av[a] = As[kk + lq][wi + 16 * a + lp];
#pragma unroll
      for (int b = 0; b < NBB; ++b) bv[b] = Bs[kk + lq][wj + 16 * b + lr];
#pragma unroll
      for (int a = 0; a < MA; ++a)
#pragma unroll
        for (int b = 0; b < NBB; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (DOT) {
      // column sums against cdot: lanes lr + 16 q hold rows of one column
#pragma unroll
      for (int b = 0; b < NBB; ++b) {
        double cs = 0.0;
#pragma unroll
        for (int a = 0; a < MA; ++a)
#pragma unroll
          for (int r = 0; r < 4; ++r) cs += acc[a][b][r] * cdl[wi + 16 * a + 4 * lq + r];
        cs += __shfl_xor(cs, 16, 64);
        cs += __shfl_xor(cs, 32, 64);
        if (lq == 0) red[wave & 3][wj + 16 * b + lr] = cs;
      }
    }
    __syncthreads();  // the panels and cdl are no longer read; red is complete
    // the next panels to LDS before this tile's stores (a wait for the
    // loads would otherwise wait for the stores too)
    if (more) {
      pb.put(Bs, 1.0);
      if (newrow) {
        Panel64<NT> pa;
        pa.fetch(g.A, g.lda, g.M, In * GT, a16);
        pa.put(As, g.alpha);
        if (DOT && tid < GT) cdl[tid] = In * GT + tid < g.M ? g.cdot[In * GT + tid] : 0.0;
      }
    }
    if (DOT && tid < GT) {
      const int j = j0 + tid;
      if (j < g.N && j >= g.jfrom)
        g.part[(int64_t)I * g.ldp + j] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    }
#pragma unroll
    for (int a = 0; a < MA; ++a)
#pragma unroll
      for (int b = 0; b < NBB; ++b) {
        const int i = i0 + wi + 16 * a + 4 * lq;
        const int j = j0 + wj + 16 * b + lr;
        if (j >= g.N || i >= g.M) continue;
        double* c = g.C + i + (int64_t)j * g.ldc;
        if (vec4 && i + 3 < g.M) {
          *reinterpret_cast<d4*>(c) = acc[a][b];
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (i + r < g.M) c[r] = acc[a][b][r];
        }
      }
    __syncthreads();  // the next panels are in LDS; red may be rewritten
    I = In;
    J = Jn;
  }
}

static int trail_cus() {
  static int cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c <= 0)
      c = 256;
    cus[dev] = c;
  }
  return cus[dev];
}

static void launch_trail64(int M, int N, double alpha, const double* A, int lda, const double* B,
                           int ldb, double beta, double* C, int ldc, const double* cdot,
                           double* part, int ldp, int jfrom, hipStream_t st) {
  TrailArgs g{M, N, alpha, beta, A, lda, B, ldb, C, ldc, cdot, part, ldp, jfrom};
  const int tm = (int)ceil_div(M, GT), tn = (int)ceil_div(N, GT);
  const int ntile = tm * tn;
  const int G = std::min(ntile, trail_cus());
  if (cdot) hipLaunchKernelGGL(k_trail64<true>, dim3(G), dim3(1024), 0, st, g, tn, ntile);
  else hipLaunchKernelGGL(k_trail64<false>, dim3(G), dim3(1024), 0, st, g, tn, ntile);
}

}  // namespace

int dgemm_trail64(int M, int N, const double* A, int lda, const double* B, int ldb, double* C,
                  int ldc, const double* cdot, double* part, int ldp, int jfrom, hipStream_t st) {
  static_assert(TRAIL_TILE == GT, "partials per 128-row tile");
  if (M <= 0 || N <= 0) return DDC_OK;
  launch_trail64(M, N, -1.0, A, lda, B, ldb, 1.0, C, ldc, cdot, part, ldp, jfrom, st);
  DDC_LAUNCHED();
  return DDC_OK;
}

int dgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
               const double* B, int ldb, double beta, double* C, int ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return DDC_OK;
  const bool TA = ta == 'T' || ta == 't', TB = tb == 'T' || tb == 't';
  if (!TA && TB && K == 64) {
    // the shape of the trailing update: its own kernel
    launch_trail64(M, N, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, nullptr, 0, 0, st);
    DDC_LAUNCHED();
    return DDC_OK;
  }
  GemmArgs g{M, N, K, alpha, beta, A, lda, B, ldb, C, ldc};
  static const int nw = getenv("DDCSVD_GEMM_WAVES") ? atoi(getenv("DDCSVD_GEMM_WAVES")) : 8;
  if (nw == 4) launch_dgemm<4>(TA, TB, g, st);
  else launch_dgemm<8>(TA, TB, g, st);
  DDC_LAUNCHED();
  return DDC_OK;
}

}  // namespace ddc
