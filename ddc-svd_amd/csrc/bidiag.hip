// bidiag.hip -- Golub-Kahan Householder bidiagonalization on gfx950.
//
// Replaces bidiag_par (bidiag_par.c:34-450) and its ten OpenCL kernels.  The
// result is the reference's: alpha, beta and the unit-norm reflectors
// (H = I - 2 v v^T, sign(0) = +1, bidiag.c:76-124) stored in place in A.
//
// Scheme (m >= n).  Panels of NB columns.  Inside a panel the trailing
// matrix is kept in its panel-start state A0 and every step uses the
// low-rank form  M_t = A0 - U Y^T - X V^T  of the partially reduced matrix
// (u_s / v_s: left / right reflectors of the panel, y_s = 2 M^T u_s,
// x_s = 2 M v_s).  A step t (global column k) is three stream-ordered
// kernels:
//
//   left_gemv   g = A0[k:m, k+1:n]^T c           2-D grid, partial sums
//   right_gemv  left finalization on each column chunk:
//                 A0^T u = (g + s|c| A0[k, :]) / rl,
//                 y_t = 2(A0^T u - Y (U^T u) - V (X^T u)); row k of M -> r;
//               then h = A0[k+1:m, k+1:n] r      2-D grid, partial sums
//   right_fin   A0 v = (h + s|r| A0[:, k+1]) / rl;
//               x_t = 2(A0 v - U (Y^T v) - X (V^T v)); column k+1 of M -> c
//
// The GEMVs multiply by the raw column c / row r of M, so their loads wait
// for nothing; the reflector u = (c + s|c| e_k)/rl (v likewise) enters in
// the fin kernels, which rebuild |c|^2 from the partial sums left by the
// producer (every work-group redoes that tiny fixed-order reduction, so no
// grid-wide synchronisation is needed and results are bitwise
// reproducible).  The two GEMVs are the only passes over the trailing
// matrix: 16 bytes per trailing element per step.  After
// NB steps the trailing matrix gets A0 -= [U X] [Y V]^T, one FP64 MFMA GEMM
// of inner dimension 2 NB (gemm.hip, one read and one write), which also
// takes the partial sums of the next panel's first left GEMV while each
// tile is in registers, so that step launches no left GEMV.
#include <algorithm>
#include <cstdlib>

#include "ddc_common.h"

namespace ddc {

constexpr int NB = 32;        // panel width
constexpr int RC = 2048;      // rows per left_gemv chunk (8 per thread)
constexpr int CBL = 8;        // columns per left_gemv work-group
constexpr int RBR = 512;      // rows per right_gemv block (2 per thread)
constexpr int CCR = 128;      // columns per right_gemv chunk
constexpr int DPW = 8;        // panel dot products per dot work-group
constexpr int FIN = 64;       // items per finalize work-group (one wave)
constexpr int DSTRIDE = 2 * NB + 1;
// bd_left_gemv's reduction array red[4][CBL] also serves its dot work-groups
static_assert(CBL >= DPW, "left_gemv: red[][CBL] must hold DPW dot products");

// TS: storage type of the matrix and of the reflectors kept in it (double,
// or float for the mixed-precision form); all vector arithmetic is FP64.
template <typename TS>
struct StepArgs {
  int m, n, lda, ldy;
  int k, p, t;
  TS* A;
  double* alpha;
  double* beta;
  double* Y;  // n x NB, ld ldy: y_s
  double* V;  // n x NB, ld ldy: v_s (contiguous copy of the row reflectors)
  double* X;  // m x NB, ld lda: x_s
  double* Ucopy;  // m x NB, ld lda: FP64 copy of u_s for the trailing update, or null
  double* cvec;  // current column of M (rows k..m-1)
  double* rvec;  // current row of M (columns k+1..n-1)
  double* Lpart; // left_gemv partials [chunk][ldy]
  double* Rpart; // right_gemv partials [chunk][lda]
  double* Ldot;  // panel dots with c [chunk][2*NB]
  double* Rdot;  // panel dots with r [chunk][DSTRIDE]
  double* cnorm; // partial |c|^2 [ncn]
  double* rnorm; // partial |r|^2 [nrn]
  int ncn, nrn;
};

// Two consecutive stored entries as doubles (p 2-element aligned).
__device__ __forceinline__ double2 ld2(const double* p) {
  return *reinterpret_cast<const double2*>(p);
}
__device__ __forceinline__ double2 ld2(const float* p) {
  const float2 v = *reinterpret_cast<const float2*>(p);
  return make_double2((double)v.x, (double)v.y);
}

// Two consecutive stored entries, raw (no conversion: the loads of a batch
// can all be in flight before the first use).
template <typename TS>
struct Raw2;
template <>
struct Raw2<double> {
  using T = double2;
};
template <>
struct Raw2<float> {
  using T = float2;
};
template <typename TS>
__device__ __forceinline__ typename Raw2<TS>::T ldraw2(const TS* p) {
  return *reinterpret_cast<const typename Raw2<TS>::T*>(p);
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

// Sum of cnt partials in a fixed order; every lane of the wave gets it.
__device__ __forceinline__ double sum_partials(const double* part, int cnt) {
  double s = 0.0;
  for (int i = (int)(threadIdx.x & 63); i < cnt; i += 64) s += part[i];
  return wave_sum(s);
}

// sum_{q<cnt} p[q*stride] in a fixed order, with the loads issued in
// batches of BATCH (clamped addresses, unused past cnt) so that their
// latencies overlap.  The order of the additions does not depend on BATCH.
template <int BATCH = 32>
__device__ __forceinline__ double sum_strided(const double* p, int cnt, int64_t stride) {
  double acc = 0.0;
  for (int q = 0; q < cnt; q += BATCH) {
    double v[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; ++u) v[u] = p[(int64_t)min(q + u, cnt - 1) * stride];
#pragma unroll
    for (int u = 0; u < BATCH; ++u)
      if (q + u < cnt) acc += v[u];
  }
  return acc;
}

// Householder data of the vector x with |x|^2 = sumsq and first entry x1,
// as the reference forms it (update_scale_matcol.cl:58-82): diag =
// -sign(x1)|x|, reflector = (x + sign(x1)|x| e1) / rl with
// rl = sqrt(2) sqrt(|x|^2 + ||x| x1|).  A zero vector gives rl = 0 and is
// treated as "no reflection" (the reference divides by zero there).
struct Refl {
  double nrm, sgn, rl;
  __device__ __forceinline__ void make(double sumsq, double x1) {
    nrm = sqrt(sumsq);
    sgn = (x1 < 0) ? -1.0 : 1.0;
    rl = sqrt(2.0) * sqrt(nrm * nrm + fabs(nrm * x1));
  }
  __device__ __forceinline__ double diag() const { return rl != 0.0 ? -sgn * nrm : 0.0; }
  __device__ __forceinline__ double vec(double x, bool first) const {
    if (rl == 0.0) return 0.0;
    return (first ? x + sgn * nrm : x) / rl;
  }
  // (w^T reflector) from (w^T x) and the first entry w1 of w
  __device__ __forceinline__ double dot(double wx, double w1) const {
    return rl != 0.0 ? (wx + sgn * nrm * w1) / rl : 0.0;
  }
};

// ---------------------------------------------------------------------------

template <typename TS>
__global__ __launch_bounds__(256) void bd_init(int m, const TS* A,
                                               double* cvec, double* cnorm) {
  __shared__ double red[4];
  int row = blockIdx.x * 256 + threadIdx.x;
  double c = (row < m) ? (double)A[row] : 0.0;
  if (row < m) cvec[row] = c;
  double s = block_sum256(c * c, red);
  if (threadIdx.x == 0) cnorm[blockIdx.x] = s;
}

// left_gemv: grid.x = ncb column blocks + dot blocks, grid.y = row chunks.
// NQ: pairs of rows per thread (rows per work-group 512 NQ).
template <typename TS, int NQ>
__global__ __launch_bounds__(256) void bd_left_gemv(StepArgs<TS> a, int ncb, int rb0) {
  __shared__ double red[4][CBL];
  __shared__ double s_scal;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = a.k;

  // The product is taken with the raw column c (rows k..m-1), not with the
  // reflector u = (c + s|c| e_k) / rl: no load waits for the norm of c, and
  // the left finalization (in bd_right_gemv, or bd_left_fin at k = n-2)
  // applies  A0^T u = (A0^T c + s|c| A0[k, :]) / rl.
  // rows of this thread: r0 + 512 q + {0, 1}, q < NQ (coalesced pairs)
  const int r0 = rb0 + blockIdx.y * (512 * NQ) + 2 * tid;
  double cv[2 * NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 512 * q + h;
      cv[2 * q + h] = (r >= k && r < a.m) ? a.cvec[r] : 0.0;
    }
  }

  // Loads without conditions, so that all CBL * NQ of them are in flight
  // together: rows past the column (r >= lda) are clamped to the last row
  // pair and columns past the matrix to its last column; their c entries
  // are zero or their sums are never stored.
  int rq[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) rq[q] = min(r0 + 512 * q, a.lda - 2);

  if ((int)blockIdx.x < ncb) {
    const int j0 = k + 1 + blockIdx.x * CBL;
    const int jn = min(CBL, a.n - j0);
    // All CBL * NQ loads are issued before the first product: without the
    // scheduling barrier the compiler interleaves them one load at a time
    // (fewer registers, but one load in flight per wave)
    typename Raw2<TS>::T v[CBL][NQ];
#pragma unroll
    for (int c = 0; c < CBL; ++c) {
      const TS* col = a.A + (int64_t)(j0 + min(c, jn - 1)) * a.lda;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v[c][q] = ldraw2(col + rq[q]);
    }
    __builtin_amdgcn_sched_barrier(0);
    double acc[CBL];
#pragma unroll
    for (int c = 0; c < CBL; ++c) {
      acc[c] = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        acc[c] += (double)v[c][q].x * cv[2 * q] + (double)v[c][q].y * cv[2 * q + 1];
    }
    // lanes 8c .. 8c+7 end with column c's wave sum
    const double sc = reduce_scatter<double, CBL>(acc);
    if ((lane & (64 / CBL - 1)) == 0) red[wave][lane / (64 / CBL)] = sc;
    __syncthreads();
    if (tid < jn)
      a.Lpart[(int64_t)blockIdx.y * a.ldy + j0 + tid] =
          red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    // the reflector itself goes to its final place, column k (not read by
    // any work-group of this launch)
    if (blockIdx.x == 0) {
      if (wave == 0) {
        double s = sum_partials(a.cnorm, a.ncn);
        if (lane == 0) s_scal = s;
      }
      __syncthreads();
      Refl R;
      R.make(s_scal, a.cvec[k]);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = r0 + 512 * q + h;
          if (r >= k && r < a.m) {
            const double u = R.vec(cv[2 * q + h], r == k);
            a.A[r + (int64_t)k * a.lda] = (TS)u;
            if (a.Ucopy) a.Ucopy[r + (int64_t)a.t * a.lda] = u;
          }
        }
      if (blockIdx.y == 0 && tid == 0) a.alpha[k] = R.diag();
    }
  } else {
    // panel dots with the raw column c: u_s^T c (s < t) and x_s^T c (s < t)
    const int q0 = (blockIdx.x - ncb) * DPW;
    const int nd = 2 * a.t;
    double acc[DPW];
#pragma unroll
    for (int e = 0; e < DPW; ++e) {
      acc[e] = 0.0;
      const int s = min(q0 + e, nd - 1);  // clamped: sums past nd are not stored
      if (s < a.t) {
        const TS* col = a.A + (int64_t)(a.p + s) * a.lda;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const double2 v = ld2(col + rq[q]);
          acc[e] += v.x * cv[2 * q] + v.y * cv[2 * q + 1];
        }
      } else {
        const double* col = a.X + (int64_t)(s - a.t) * a.lda;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const double2 v = ld2(col + rq[q]);
          acc[e] += v.x * cv[2 * q] + v.y * cv[2 * q + 1];
        }
      }
    }
    const double se = reduce_scatter<double, DPW>(acc);
    if ((lane & (64 / DPW - 1)) == 0) red[wave][lane / (64 / DPW)] = se;
    __syncthreads();
    if (tid < DPW && q0 + tid < nd)
      a.Ldot[(int64_t)blockIdx.y * (2 * NB) + q0 + tid] =
          red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
  }
}

// The left GEMV of the first column k of a panel when the trailing update
// has already formed its partial sums, one per 128-row tile (nt of them, in
// Lpart rows 0..nt-1): blocks [0, ncs) sum them in tile order into Lpart
// row 0 for the columns k+1..n-1; the other blocks store the reflector of
// column k (rows k..m-1) in A and Ucopy and its diagonal, as bd_left_gemv
// does.
template <typename TS>
__global__ __launch_bounds__(256) void bd_left_red(StepArgs<TS> a, int nt, int ncs) {
  __shared__ double s_scal;
  const int tid = threadIdx.x;
  const int k = a.k;
  if ((int)blockIdx.x < ncs) {
    const int j = k + 1 + blockIdx.x * 256 + tid;
    if (j < a.n) {
      const double g = sum_strided<32>(a.Lpart + j, nt, a.ldy);
      a.Lpart[j] = g;
    }
    return;
  }
  if (tid < 64) {
    const double s = sum_partials(a.cnorm, a.ncn);
    if (tid == 0) s_scal = s;
  }
  __syncthreads();
  Refl R;
  R.make(s_scal, a.cvec[k]);
  const int r = k + ((int)blockIdx.x - ncs) * 256 + tid;
  if (r < a.m) {
    const double u = R.vec(a.cvec[r], r == k);
    a.A[r + (int64_t)k * a.lda] = (TS)u;
    if (a.Ucopy) a.Ucopy[r + (int64_t)a.t * a.lda] = u;
  }
  if ((int)blockIdx.x == ncs && tid == 0) a.alpha[k] = R.diag();
}

// left_fin: one thread per trailing column j in [k+1, n), one wave per group.
// Launched only at k = n-2 (no right GEMV follows); every other step does
// this work in bd_right_gemv's prologue.
template <typename TS>
__global__ __launch_bounds__(FIN) void bd_left_fin(StepArgs<TS> a, int nchL) {
  __shared__ double dd[2 * NB];  // u_s^T u, x_s^T u
  __shared__ double uk[NB + 1];  // u_s[k], s <= t
  __shared__ double xk[NB];      // x_s[k], s < t
  const int tid = threadIdx.x;
  const int k = a.k, t = a.t;
  const int j = k + 1 + blockIdx.x * FIN + tid;
  const bool jv = j < a.n;
  const int jc = jv ? j : k + 1;  // a valid column for the loads of idle lanes

  // Every load first: none depends on the reflector, so their latencies
  // overlap instead of following the block's setup.
  double ys[NB], vs[NB];
  const double* Yj = a.Y + jc;
  const double* Vj = a.V + jc;
#pragma unroll
  for (int g = 0; g < NB / 8; ++g) {
    if (g * 8 < t) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int s = min(g * 8 + u, t - 1);  // clamped: in bounds, unused past t
        ys[g * 8 + u] = Yj[(int64_t)s * a.ldy];
        vs[g * 8 + u] = Vj[(int64_t)s * a.ldy];
      }
    }
  }
  double r = (double)a.A[k + (int64_t)jc * a.lda];
  double cn = 0.0;
  for (int i = tid; i < a.ncn; i += FIN) cn += a.cnorm[i];
  const double ck = a.cvec[k];
  const bool dotl = tid < 2 * t;
  const int td = dotl ? tid : 0;
  const double w1 = (td < t) ? (double)a.A[k + (int64_t)(a.p + td) * a.lda]
                             : a.X[k + (int64_t)(td - t) * a.lda];
  const double ds = sum_strided(a.Ldot + td, nchL, 2 * NB);
  const double g = sum_strided(a.Lpart + jc, nchL, a.ldy);  // A0[k:m, j]^T c

  Refl R;
  R.make(wave_sum(cn), ck);
  if (dotl) dd[tid] = R.dot(ds, w1);
  double y = R.dot(g, r);  // A0[k:m, j]^T u, with r = A0[k, j]
  if (tid < t) {
    uk[tid] = (double)a.A[k + (int64_t)(a.p + tid) * a.lda];
    xk[tid] = a.X[k + (int64_t)tid * a.lda];
  }
  if (tid == t) uk[t] = R.vec(ck, true);
  __syncthreads();

  double rr = 0.0;
  if (jv) {
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      if (s < t) {
        y -= ys[s] * dd[s] + vs[s] * dd[t + s];
        r -= uk[s] * ys[s] + xk[s] * vs[s];
      }
    }
    y *= 2.0;
    a.Y[j + (int64_t)t * a.ldy] = y;
    r -= uk[t] * y;
    a.rvec[j] = r;
    rr = r * r;
  }
  double s = wave_sum(rr);
  if (tid == 0) a.rnorm[blockIdx.x] = s;
}

// right_gemv with the left finalization fused in.  grid.x = nrb row blocks
// + 1 owner block, grid.y = column chunks of CC = 128 or 64 columns (the
// narrower chunks for the smaller trailing matrices, where twice the
// work-groups keep more loads in flight).
//
// Every work-group first forms row k of M on its chunk, as bd_left_fin does
// for all columns (bitwise the same values in every work-group: same code,
// same order).  With P = Y (U^T c) + V (X^T c) and Q = Y U[k, :]^T +
// V X[k, :]^T (panel terms, j-th entries) the left finalization is
//   y_t = 2 ((g - P) + s|c| (A0[k, :] - Q)) / rl,   r = A0[k, :] - Q - u_t[k] y_t,
// so the column threads need the reflector data only at the very end.  The
// row blocks then stream A0[rows, chunk] r; the owner block stores y_t and r
// and leaves the chunk's partial sums of |r|^2, y_s^T r (s <= t) and
// v_s^T r (s < t) for right_fin.  This saves the launch of bd_left_fin.
template <typename TS, int CC, int PF>
__global__ __launch_bounds__(256) void bd_right_gemv(StepArgs<TS> a, int nrb, int rb0, int nchL) {
  static_assert(CC == 128 || CC == 64, "a whole number of waves per part of the chunk");
  static_assert(PF % 8 == 0 && PF >= 8, "columns prefetched in batches of eight");
  constexpr int TPC = 256 / CC;         // threads per column: the column's parts
  constexpr int H = NB / TPC;           // panel columns per part
  constexpr int WPP = CC / 64;          // waves per part
  __shared__ double rs[CC];             // r on the chunk
  __shared__ double yv[CC];             // y_t on the chunk
  __shared__ double pq[TPC - 1][2][CC]; // the other parts' P, Q
  __shared__ double ds[2 * NB];         // u_s^T c, x_s^T c (s < t)
  __shared__ double uk[NB + 1];         // U[k, s], s <= t (u_t[k] from the reflector)
  __shared__ double xk[NB];             // X[k, s], s < t
  __shared__ double rfl[3];             // the left reflector: |c|, sign, rl
  __shared__ double red[4][32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = a.k, t = a.t;
  // (row block bx, column chunk by), XCD-aware: work-groups go to the eight
  // XCDs round-robin by linear id, so the ids are permuted to give each XCD
  // whole runs of consecutive (bx, by); the row blocks of a chunk then share
  // one L2 for the chunk's panel values (Y, V, the left GEMV's partials)
  // instead of each XCD fetching its own copy.
  int bx, by;
  {
    const int gx = (int)gridDim.x, total = gx * (int)gridDim.y;
    const int lin = (int)blockIdx.x + (int)blockIdx.y * gx;
    const int q = total / 8;
    const int L = lin < 8 * q ? (lin % 8) * q + lin / 8 : lin;
    bx = L % gx;
    by = L / gx;
  }
  const int j0 = k + 1 + by * CC;
  const int jn = min(CC, a.n - j0);
  const int col = tid & (CC - 1), part = tid / CC;
  const bool jv = col < jn;
  const int jc = jv ? j0 + col : k + 1;  // a valid column for the loads of idle lanes

  // A row block (rows r0, r0 + 1 of each thread) issues its first PF
  // columns of A0 before the prologue (rows past the column clamped; unused),
  // so that they stream during it.
  const bool rowblk = bx < nrb;
  const int r0 = rb0 + bx * RBR + 2 * tid;
  const TS* base = a.A + min(r0, a.lda - 2) + (int64_t)j0 * a.lda;
  double2 v0[PF / 8][8];
  if (rowblk) {
#pragma unroll
    for (int i = 0; i < PF / 8; ++i)
#pragma unroll
      for (int u = 0; u < 8; ++u) v0[i][u] = ld2(base + (int64_t)min(8 * i + u, jn - 1) * a.lda);
  }

  // Every load first.  Thread (col, part) holds Y[j, s], V[j, s] for the
  // panel columns s = part H + i < t.
  double ys[H] = {}, vs[H] = {};
  const double* Yj = a.Y + jc + (int64_t)(part * H) * a.ldy;
  const double* Vj = a.V + jc + (int64_t)(part * H) * a.ldy;
  const int th = t - part * H;  // panel columns of this part
#pragma unroll
  for (int g = 0; g < H / 8; ++g) {
    if (g * 8 < th) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = min(g * 8 + u, th - 1);  // clamped: in bounds, unused past th
        ys[g * 8 + u] = Yj[(int64_t)i * a.ldy];
        vs[g * 8 + u] = Vj[(int64_t)i * a.ldy];
      }
    }
  }
  double gsum = 0.0, rk = 0.0;
  if (part == 0) {
    rk = (double)a.A[k + (int64_t)jc * a.lda];
    gsum = sum_strided<8>(a.Lpart + jc, nchL, a.ldy);  // A0[k:m, j]^T c
  }
  if (wave == 2) {
    // the left reflector (from the partial sums of |c|^2) and u_t[k]
    double cn = 0.0;
    for (int i = lane; i < a.ncn; i += 64) cn += a.cnorm[i];
    const double ck = a.cvec[k];
    Refl R;
    R.make(wave_sum(cn), ck);
    if (lane == 0) {
      rfl[0] = R.nrm;
      rfl[1] = R.sgn;
      rfl[2] = R.rl;
      uk[t] = R.vec(ck, true);
    }
  } else if (wave == 3) {
    const bool dl = lane < 2 * t;
    const double d = sum_strided<8>(a.Ldot + (dl ? lane : 0), nchL, 2 * NB);
    if (dl) ds[lane] = d;
    if (lane < t) {
      uk[lane] = (double)a.A[k + (int64_t)(a.p + lane) * a.lda];
      xk[lane] = a.X[k + (int64_t)lane * a.lda];
    }
  }
  __syncthreads();

  // panel terms of this part
  double P = 0.0, Q = 0.0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    if (i < th) {
      const int sx = part * H + i;
      P += ys[i] * ds[sx] + vs[i] * ds[t + sx];
      Q += ys[i] * uk[sx] + vs[i] * xk[sx];
    }
  }
  if (part > 0) {
    pq[part - 1][0][col] = P;
    pq[part - 1][1][col] = Q;
  }
  __syncthreads();
  double y = 0.0, r = 0.0;
  if (part == 0) {
    Refl R;
    R.nrm = rfl[0];
    R.sgn = rfl[1];
    R.rl = rfl[2];
#pragma unroll
    for (int q = 0; q < TPC - 1; ++q) {
      P += pq[q][0][col];
      Q += pq[q][1][col];
    }
    y = 2.0 * R.dot(gsum - P, rk - Q);
    r = rk - Q - uk[t] * y;
    if (!jv) y = r = 0.0;
    rs[col] = r;
    yv[col] = y;
  }
  __syncthreads();

  if (rowblk) {
    // Two batches of eight columns in flight: the next batch is issued
    // before the current one is consumed (one batch at a time left each
    // wave idle for a full memory latency per batch; 8192^2: 283 -> 278 ms
    // per SVD in this kernel).  Sixteen loads issued at once and then
    // consumed, or two row pairs per thread, were slower (296 and 301 ms).
    double h0 = 0.0, h1 = 0.0;
    double2 va[8], vb[8];
    auto load8 = [&](double2 (&v)[8], int c0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = ld2(base + (int64_t)(c0 + u) * a.lda);
    };
    auto use8 = [&](const double2 (&v)[8], int c0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        h0 += v[u].x * rs[c0 + u];
        h1 += v[u].y * rs[c0 + u];
      }
    };
    int c = 0;
    if (jn >= PF) {
      // the prefetched batches are consumed before the loop, so that the
      // loop is entered with one batch (va) in flight, as it is repeated
      const bool more = jn >= PF + 8;
      if (more) load8(va, PF);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < PF / 8; ++i) use8(v0[i], 8 * i);
      c = PF;
      if (more) {
        // at the top of the loop va holds (or is loading) columns c - 8 .. c - 1
        c = PF + 8;
        while (true) {
          if (c + 8 > jn) {
            use8(va, c - 8);
            break;
          }
          load8(vb, c);
          __builtin_amdgcn_sched_barrier(0);
          use8(va, c - 8);
          c += 8;
          if (c + 8 > jn) {
            use8(vb, c - 8);
            break;
          }
          load8(va, c);
          __builtin_amdgcn_sched_barrier(0);
          use8(vb, c - 8);
          c += 8;
        }
      }
    }
    for (; c < jn; ++c) {
      const double2 v = ld2(base + (int64_t)c * a.lda);
      h0 += v.x * rs[c];
      h1 += v.y * rs[c];
    }
    double* out = a.Rpart + (int64_t)by * a.lda;
    if (r0 >= k + 1 && r0 < a.m) out[r0] = h0;
    if (r0 + 1 >= k + 1 && r0 + 1 < a.m) out[r0 + 1] = h1;
  } else {
    // owner: y_t and r, then the chunk's partial sums for right_fin in the
    // layout of Rdot: [s] y_s^T r (s <= t), [t + 1 + s] v_s^T r (s < t),
    // and |r|^2 in rnorm.  Slots of the reduction: 0..31 y_s^T r (slot t:
    // y_t^T r), 32..62 v_s^T r, 63 |r|^2; this part holds slots
    // part H + i and 32 + part H + i, i < H.
    if (part == 0 && jv) {
      a.Y[j0 + col + (int64_t)t * a.ldy] = y;
      a.rvec[j0 + col] = r;
    }
    const double rr = rs[col], yy = yv[col];
    double g[2 * H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const int sx = part * H + i;
      g[i] = (i < th) ? ys[i] * rr : (sx == t ? yy * rr : 0.0);
      g[H + i] = (i < th) ? vs[i] * rr : (sx == NB - 1 ? rr * rr : 0.0);
    }
    constexpr int LPS = 64 / (2 * H);  // lanes per slot after the reduce-scatter
    const double tot = reduce_scatter<double, 2 * H>(g);  // lane l: slot index l / LPS
    if ((lane & (LPS - 1)) == 0) red[wave][lane / LPS] = tot;
    __syncthreads();
    if (tid < 64) {
      const int pt = tid / (2 * H), i = tid % (2 * H);  // part, index within its 2H slots
      double v = red[pt * WPP][i];
#pragma unroll
      for (int w = 1; w < WPP; ++w) v += red[pt * WPP + w][i];
      const bool ysl = i < H;
      const int sx = pt * H + (ysl ? i : i - H);
      double* Rd = a.Rdot + (int64_t)by * DSTRIDE;
      if (ysl) {
        if (sx <= t) Rd[sx] = v;
      } else if (sx < t) {
        Rd[t + 1 + sx] = v;
      } else if (sx == NB - 1) {
        a.rnorm[by] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// One pass over the trailing matrix per column step (m - k <= FG_ROWS).
//
// The two GEMVs of a step read the same block A0[k:m, k+1:n]: the left one
// for g_j = A0[:, j]^T c, the right one for h = A0 r, whose r_j needs only
// g_j and column j's panel terms.  A work-group that holds a whole column
// can therefore finish both with one read (the BLAS 2.5 reorganisation of
// Howell, Demmel, Fulton, Hammarling and Marmol, ACM TOMS 34(3), 2008):
// per column it reduces g_j over all rows, forms y_t[j] and r_j as
// bd_right_gemv's prologue does, and adds A0[:, j] r_j to its rows' partial
// h.  Work-group w takes the columns k+1+w, k+1+w+G, ...; each column
// stays in registers (RPT rows per thread) from its load to its last use,
// and the next D-1 columns load meanwhile (a ring of D columns, indexed
// statically by unrolling the column loop D times).  Two barriers per
// column: the per-wave partial g is complete; r_j is known.  Work-group
// w's partial h (all rows) goes to Rpart row w, its partial sums over its
// columns of |r|^2, y_s^T r and v_s^T r to Rdot row w and rnorm[w], as
// chunk w of the right GEMV would.  The panel dots u_s^T c and x_s^T c
// (s < t) arrive as the previous bd_right_fin's per-group partials (nldot
// rows of Ldot); the reflector of column k is stored by slices of rows
// across the groups.
constexpr int FG_T = 512;               // threads per work-group (256 registers each)
constexpr int FG_ROWS = FG_T * 16;      // rows k..m-1 held: m - k <= 8192
constexpr int FG_GMAX = 256;            // work-groups (one per CU)
constexpr int FG_MAXCOL = 512;          // columns per work-group: n - k - 1 <= 256 x 512

// Buffer loads: word 3 of a raw buffer resource on gfx9 targets (32-bit
// data format, no swizzle); out-of-range offsets read zero.
constexpr int BUF_RSRC_WORD3 = 0x00020000;
template <typename TS>
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, int voff);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int voff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}
template <>
__device__ __forceinline__ double buf_load<float>(__amdgpu_buffer_rsrc_t r, int voff) {
  return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}

// Wave sum by DPP (row-local butterflies, then the row broadcasts of gfx9),
// in a fixed order; every lane gets lane 63's total.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_take(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// (every lane written: no old value to set)
template <int CTRL>
__device__ __forceinline__ double dpp_all(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane63(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_all<0xB1>(v);         // quad_perm [1,0,3,2]: pairs
  v += dpp_all<0x4E>(v);         // quad_perm [2,3,0,1]: quads
  v += dpp_all<0x141>(v);        // row_half_mirror: eights
  v += dpp_all<0x140>(v);        // row_mirror: rows of 16
  v += dpp_take<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_take<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3: lane 63 has the total
  return lane63(v);
}
// Three wave totals with their stages interleaved (independent chains).
__device__ __forceinline__ void wave_total3(double& a, double& b, double& c) {
  a += dpp_all<0xB1>(a);
  b += dpp_all<0xB1>(b);
  c += dpp_all<0xB1>(c);
  a += dpp_all<0x4E>(a);
  b += dpp_all<0x4E>(b);
  c += dpp_all<0x4E>(c);
  a += dpp_all<0x141>(a);
  b += dpp_all<0x141>(b);
  c += dpp_all<0x141>(c);
  a += dpp_all<0x140>(a);
  b += dpp_all<0x140>(b);
  c += dpp_all<0x140>(c);
  a += dpp_take<0x142, 0xA>(a);
  b += dpp_take<0x142, 0xA>(b);
  c += dpp_take<0x142, 0xA>(c);
  a += dpp_take<0x143, 0xC>(a);
  b += dpp_take<0x143, 0xC>(b);
  c += dpp_take<0x143, 0xC>(c);
  a = lane63(a);
  b = lane63(b);
  c = lane63(c);
}

// num / d from ir = 1/d (0 for d = 0, giving 0): q0 = num ir, then one
// correction q0 + (num - q0 d) ir with fused multiply-adds.
__device__ __forceinline__ double rl_quot(double num, double d, double ir) {
  const double q0 = num * ir;
  return fma(fma(-q0, d, num), ir, q0);
}

template <typename TS, int RPT, int D, bool PAIR>
__global__ __launch_bounds__(FG_T, 1) void bd_fused_gemv(StepArgs<TS> a, int nldot) {
  constexpr int NW = FG_T / 64;
  // per column (pair) parity: partial g by wave, A0[k, j], P, Q
  __shared__ double red[2][2 * NW + 6];
  __shared__ double dsp[NW][2 * NB];  // partial panel dots
  __shared__ double cvl[FG_T * RPT];  // c on rows k..k + FG_T RPT - 1
  __shared__ double yst[FG_MAXCOL];   // y_t[j], r_j of the group's columns (stored
  __shared__ double rst[FG_MAXCOL];   // after the loop: no store among its loads)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = a.k, t = a.t, m = a.m, n = a.n;
  const int G = (int)gridDim.x, w = (int)blockIdx.x;
  const int ncol = (n - k - 1 - w + G - 1) / G;  // columns of this group (>= 1)

  // Per-step values first, every load unconditional and straight-line (a
  // load under a branch, or a wait inside a loop, made the compiler wait
  // for every load in flight): c on this thread's rows, the partial norms
  // of c, the panel dots' partial rows, U[k, :] and X[k, :].  Then the
  // first D-1 columns.
  double cv[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) cv[q] = a.cvec[min(k + tid + FG_T * q, m - 1)];
  constexpr int NCN = 2, NLD = 16;  // one batch; the rest (wide shapes) in loops
  double cnb[NCN], ldb[NLD];
#pragma unroll
  for (int u = 0; u < NCN; ++u) cnb[u] = a.cnorm[min(lane + 64 * u, a.ncn - 1)];
  const int dl = lane < 2 * t ? lane : 0;
  if (lane < 2 * t) {  // the lanes that use them only
#pragma unroll
    for (int u = 0; u < NLD; ++u)
      ldb[u] = a.Ldot[(int64_t)min(wave + NW * u, max(nldot - 1, 0)) * (2 * NB) + dl];
  } else {
#pragma unroll
    for (int u = 0; u < NLD; ++u) ldb[u] = 0.0;
  }
  const int sl = min(lane, max(t - 1, 0));
  const int slw = wave == 0 ? sl : 0;
  const double ukr = (double)a.A[k + (int64_t)(a.p + sl) * a.lda];
  const double xkr = a.X[k + (int64_t)sl * a.lda];
  const double ck = a.cvec[k];

  // Column i of this group through buffer loads: the resource (scalar)
  // ends at row m, so rows past it read zero, and is empty for i >= ncol.
  // Every load is issued unconditionally, so that the compiler can wait
  // for one ring slot while the later ones stay in flight.
  const int voff = (k + tid) * (int)sizeof(TS);
  auto fetch = [&](int i, double (&v)[RPT], double& yv, double& vv) {
    const bool ok = i < ncol;
    const int j = ok ? k + 1 + w + G * i : k + 1;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<TS*>(a.A + (int64_t)j * a.lda), 0, ok ? m * (int)sizeof(TS) : 0,
        BUF_RSRC_WORD3);
#pragma unroll
    for (int q = 0; q < RPT; ++q) v[q] = buf_load<TS>(rs, voff + q * FG_T * (int)sizeof(TS));
    // Y[j, s], V[j, s] on lane s < t of wave 0 (the other waves load one
    // line, unused: every load stays unconditional)
    yv = a.Y[j + (int64_t)slw * a.ldy];
    vv = a.V[j + (int64_t)slw * a.ldy];
  };

  // columns in flight ahead of the current one (pairs: the current pair)
  constexpr int AHEAD = PAIR ? D - 2 : D - 1;
  double ring[D][RPT];
  double ryv[D], rvv[D];
#pragma unroll
  for (int u = 0; u < AHEAD; ++u) fetch(u, ring[u], ryv[u], rvv[u]);

#pragma unroll
  for (int q = 0; q < RPT; ++q) cvl[tid + FG_T * q] = k + tid + FG_T * q < m ? cv[q] : 0.0;
  // the reflector of c, formed by every wave alike (same partials, same order)
  double cn = 0.0;
#pragma unroll
  for (int u = 0; u < NCN; ++u)
    if (lane + 64 * u < a.ncn) cn += cnb[u];
  for (int i = lane + 64 * NCN; i < a.ncn; i += 64) cn += a.cnorm[i];
  Refl R;
  R.make(wave_total(cn), ck);
  // partial panel dots: lane = dot index, wave = residue of the partial row
  double d = 0.0;
#pragma unroll
  for (int u = 0; u < NLD; ++u)
    if (wave + NW * u < nldot) d += ldb[u];
  for (int q = wave + NW * NLD; q < nldot; q += NW) d += a.Ldot[(int64_t)q * (2 * NB) + dl];
  dsp[wave][lane] = lane < 2 * t ? d : 0.0;
  __syncthreads();  // dsp, cvl
  // every wave's per-lane panel values (lane s < t)
  double dsu = 0.0, dsx = 0.0;
  if (lane < t) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      dsu += dsp[q][lane];
      dsx += dsp[q][t + lane];
    }
  }
  const double ukl = lane < t ? ukr : 0.0, xkl = lane < t ? xkr : 0.0;
  const double ukt = R.vec(ck, true);
  // y_t[j] = 2 ((g - P) + s|c| (rk - Q)) / rl, as Refl::dot, by the
  // reciprocal of rl and one FMA residual correction (Markstein): off the
  // latency of a full division per column, with its rounding
  const double snrm = R.sgn * R.nrm, irl = R.rl != 0.0 ? 1.0 / R.rl : 0.0;

  double h[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) h[q] = 0.0;
  double oy = 0.0, ov = 0.0;  // wave 0: owner partial sums of lane s
  if constexpr (PAIR) {
  // columns in pairs: one barrier and interleaved reductions per pair
  for (int i0 = 0; i0 < ncol; i0 += D) {
#pragma unroll
    for (int u = 0; u < D; u += 2) {
      const int i = i0 + u;
      if (i < ncol) {
        const bool two = i + 1 < ncol;
        // the columns D-2 and D-1 ahead into the slots freed by i-2, i-1
        fetch(i + D - 2, ring[(u + D - 2) % D], ryv[(u + D - 2) % D], rvv[(u + D - 2) % D]);
        fetch(i + D - 1, ring[(u + D - 1) % D], ryv[(u + D - 1) % D], rvv[(u + D - 1) % D]);
        const double (&v0)[RPT] = ring[u];
        const double (&v1)[RPT] = ring[(u + 1) % D];
        // (two partial sums per column: shorter dependent chains)
        double ga[2] = {0.0, 0.0}, gb[2] = {0.0, 0.0};
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const double c = cvl[tid + FG_T * q];
          ga[q & 1] += v0[q] * c;
          gb[q & 1] += v1[q] * c;
        }
        double g0 = ga[0] + ga[1], g1 = gb[0] + gb[1];
        // wave 0 forms the panel terms of both columns (lanes s >= t add
        // zero); every wave then forms r_j itself
        const int pb = (i >> 1) & 1;
        if (wave == 0) {
          const double y0 = lane < t ? ryv[u] : 0.0, w0 = lane < t ? rvv[u] : 0.0;
          const double y1 = lane < t ? ryv[(u + 1) % D] : 0.0;
          const double w1 = lane < t ? rvv[(u + 1) % D] : 0.0;
          double P0 = y0 * dsu + w0 * dsx, Q0 = y0 * ukl + w0 * xkl;
          double P1 = y1 * dsu + w1 * dsx, Q1 = y1 * ukl + w1 * xkl;
          wave_total3(g0, P0, Q0);
          wave_total3(g1, P1, Q1);
          if (lane == 0) {
            red[pb][2 * NW + 2] = P0;
            red[pb][2 * NW + 3] = Q0;
            red[pb][2 * NW + 4] = P1;
            red[pb][2 * NW + 5] = Q1;
          }
        } else {
          double z = 0.0;
          wave_total3(g0, g1, z);
        }
        if (lane == 0) {
          red[pb][wave] = g0;
          red[pb][NW + wave] = g1;
        }
        if (tid == 0) {
          red[pb][2 * NW] = v0[0];  // row k
          red[pb][2 * NW + 1] = v1[0];
        }
        __syncthreads();
        double y[2], r[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          double gq[NW / 2];
#pragma unroll
          for (int q = 0; q < NW / 2; ++q)
            gq[q] = red[pb][c * NW + 2 * q] + red[pb][c * NW + 2 * q + 1];
#pragma unroll
          for (int h2 = NW / 4; h2 >= 1; h2 /= 2)
#pragma unroll
            for (int q = 0; q < h2; ++q) gq[q] += gq[q + h2];
          const double rk = red[pb][2 * NW + c];
          const double P = red[pb][2 * NW + 2 + 2 * c], Q = red[pb][2 * NW + 3 + 2 * c];
          y[c] = 2.0 * rl_quot((gq[0] - P) + snrm * (rk - Q), R.rl, irl);
          r[c] = rk - Q - ukt * y[c];
        }
        if (!two) y[1] = r[1] = 0.0;  // past the group's last column
        if (wave == 0) {
          if (lane == 0) {
            yst[i] = y[0];
            rst[i] = r[0];
            if (two) {
              yst[i + 1] = y[1];
              rst[i + 1] = r[1];
            }
          }
          // owner slots of lane s: y_s^T r (s < t), y_t^T r (s = t),
          // |r|^2 (s = 63); v_s^T r (s < t)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const double ys = c == 0 ? ryv[u] : ryv[(u + 1) % D];
            const double vs = c == 0 ? rvv[u] : rvv[(u + 1) % D];
            oy += (lane < t ? ys : (lane == t ? y[c] : (lane == 63 ? r[c] : 0.0))) * r[c];
            ov += (lane < t ? vs : 0.0) * r[c];
          }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) h[q] += v0[q] * r[0] + v1[q] * r[1];
      }
    }
  }
  } else {
  for (int i0 = 0; i0 < ncol; i0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int i = i0 + u;
      if (i < ncol) {
        // the column D-1 ahead into the slot freed by column i-1
        fetch(i + D - 1, ring[(u + D - 1) % D], ryv[(u + D - 1) % D], rvv[(u + D - 1) % D]);
        const double (&v)[RPT] = ring[u];
        // (four partial sums: a shorter dependent chain)
        double g4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < RPT; ++q) g4[q & 3] += v[q] * cvl[tid + FG_T * q];
        double gp = (g4[0] + g4[1]) + (g4[2] + g4[3]);
        // wave 0 forms the panel terms of column j (lanes s >= t add zero);
        // every wave then forms r_j itself: one barrier per column, with
        // the partials double-buffered by column parity
        const double yv = lane < t ? ryv[u] : 0.0, vv = lane < t ? rvv[u] : 0.0;
        const int pb = i & 1;
        if (wave == 0) {
          double P = yv * dsu + vv * dsx, Q = yv * ukl + vv * xkl;
          wave_total3(gp, P, Q);
          if (lane == 0) {
            red[pb][NW + 1] = P;
            red[pb][NW + 2] = Q;
          }
        } else {
          gp = wave_total(gp);
        }
        if (lane == 0) red[pb][wave] = gp;
        if (tid == 0) red[pb][NW] = v[0];  // row k
        __syncthreads();
        const double P = red[pb][NW + 1], Q = red[pb][NW + 2];
        double gq[NW / 2];
#pragma unroll
        for (int q = 0; q < NW / 2; ++q) gq[q] = red[pb][2 * q] + red[pb][2 * q + 1];
#pragma unroll
        for (int h2 = NW / 4; h2 >= 1; h2 /= 2)
#pragma unroll
          for (int q = 0; q < h2; ++q) gq[q] += gq[q + h2];
        const double g = gq[0];
        const double rk = red[pb][NW];
        const double y = 2.0 * rl_quot((g - P) + snrm * (rk - Q), R.rl, irl);
        const double r = rk - Q - ukt * y;
        if (wave == 0) {
          if (lane == 0) {
            yst[i] = y;
            rst[i] = r;
          }
          // owner slots of lane s: y_s^T r (s < t), y_t^T r (s = t),
          // |r|^2 (s = 63); v_s^T r (s < t)
          oy += (lane < t ? ryv[u] : (lane == t ? y : (lane == 63 ? r : 0.0))) * r;
          ov += (lane < t ? rvv[u] : 0.0) * r;
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) h[q] += v[q] * r;
      }
    }
  }
  }
  __syncthreads();  // yst, rst
  // y_t and r on the group's columns
  for (int i = tid; i < ncol; i += FG_T) {
    const int j = k + 1 + w + G * i;
    a.Y[j + (int64_t)t * a.ldy] = yst[i];
    a.rvec[j] = rst[i];
  }
  // the reflector of column k: rows k + w S .. in slices, c from LDS
  // (rows k..m-1 all held: m - k <= FG_T RPT)
  {
    const int S = (m - k + G - 1) / G;
    for (int i = k + w * S + tid; i < min(m, k + (w + 1) * S); i += FG_T) {
      const double u = R.vec(cvl[i - k], i == k);
      a.A[i + (int64_t)k * a.lda] = (TS)u;
      if (a.Ucopy) a.Ucopy[i + (int64_t)t * a.lda] = u;
    }
    if (w == 0 && tid == 0) a.alpha[k] = R.diag();
  }
  // partial h on rows k+1..m-1
  double* out = a.Rpart + (int64_t)w * a.lda;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int i = k + tid + FG_T * q;
    if (i >= k + 1 && i < m) out[i] = h[q];
  }
  if (wave == 0) {
    double* Rd = a.Rdot + (int64_t)w * DSTRIDE;
    if (lane <= t) Rd[lane] = oy;
    if (lane < t) Rd[t + 1 + lane] = ov;
    if (lane == 63) a.rnorm[w] = oy;
  }
}

// Rows per thread of the one-pass GEMV for m - k rows, and its launch.
template <typename TS>
static void launch_fused_gemv(const StepArgs<TS>& a, int G, int nldot, bool clock,
                              hipStream_t st) {
  const int rows = a.m - a.k;
  hipEvent_t ke0, ke1;
  const bool ck = clock && kclock_on() && kclock_pair(KCLOCK_FUSED_GEMV, &ke0, &ke1) == 0;
  auto go = [&](auto kern) {
    if (ck)
      hipExtLaunchKernelGGL(kern, dim3(G), dim3(FG_T), 0, st, ke0, ke1, 0, a, nldot);
    else
      hipLaunchKernelGGL(kern, dim3(G), dim3(FG_T), 0, st, a, nldot);
  };
  // ring depths: as many columns as the 256 registers hold without
  // spills; columns in pairs (an even depth) where the registers allow
  if (rows > FG_T * 8) go(bd_fused_gemv<TS, 16, 3, false>);
  else if (rows > FG_T * 4) go(bd_fused_gemv<TS, 8, 8, true>);
  else if (rows > FG_T * 2) go(bd_fused_gemv<TS, 4, 8, true>);
  else go(bd_fused_gemv<TS, 2, 8, true>);
}

// right_fin: one row in [k+1, m) per lane, FIN rows per group of NWF (4
// or 8) waves; lane i of the grid also stores entry k+1+i of the right
// reflector (the grid covers max(m, n) - k - 1 rows).  The waves share the
// loads of a row: wave w takes the panel columns s = PW w .. PW w + PW-1
// and the chunk partials q = w, w + NWF, ..., so that every load of the
// step is issued at once (one wave per group took 100 loads per lane and a
// second round trip above 32 chunks); the partial sums meet in LDS.
// noright: the row of M has a single entry, taken as beta with no
// reflection (bidiag.c:125-129).
//
// dots: also leave, per group, the next step's panel dots over its rows,
// u_s^T c and x_s^T c for s <= t (c the new column), in Ldot row
// blockIdx.x at [s] and [t + 1 + s], for bd_fused_gemv.  PB: chunk
// partials per wave and batch (32 for the fused GEMV's up to 256).
template <typename TS, int PB, int NWF>
__global__ __launch_bounds__(NWF * FIN) void bd_right_fin(StepArgs<TS> a, int nccR, int noright,
                                                          int dots) {
  constexpr int PW = NB / NWF;  // panel columns per wave
  __shared__ double dyv[DSTRIDE];  // y_s^T v (s <= t), v_s^T v (s < t)
  __shared__ double yk1[NB];       // Y[k+1, s], s <= t
  __shared__ double vk1[NB];       // V[k+1, s], s < t; [t]: v[k+1]
  __shared__ double hp[NWF][FIN];  // per wave: partial A0[row, k+1:n] r
  __shared__ double dp[NWF][FIN];  // per wave: partial panel dots with r
  __shared__ double xp[NWF][FIN];  // per wave: panel terms of x
  __shared__ double cp[NWF][FIN];  // per wave: panel terms of c
  __shared__ double cnl[FIN];      // dots: the new c on the group's rows
  __shared__ double rfl[3];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = a.k, t = a.t;
  const int row = k + 1 + blockIdx.x * FIN + lane;
  const bool rv = row < a.m;
  const int rc = rv ? row : k + 1;  // a valid row for the loads of idle lanes
  const bool next = k + 1 < a.n;

  // Every load first (see bd_left_fin).
  const TS* Ur = a.A + rc + (int64_t)a.p * a.lda;
  const double* Xr = a.X + rc;
  const int s0 = PW * wave;
  double us[PW], xs[PW];
  if (s0 < t) {
#pragma unroll
    for (int i = 0; i < PW; ++i) {
      const int s = min(s0 + i, t - 1);  // clamped: in bounds, unused past t
      us[i] = (double)Ur[(int64_t)s * a.lda];
      xs[i] = Xr[(int64_t)s * a.lda];
    }
  }
  const bool dotr = !noright && lane < 2 * t + 1;
  const int td = dotr ? lane : 0;
  const int ncl = noright ? 1 : nccR;
  // (the dot loads on the 2t+1 lanes that use them only: idle quarter-waves
  // return no data)
  double pd[PB], ph[PB];
  if (dotr) {
#pragma unroll
    for (int u = 0; u < PB; ++u)
      pd[u] = a.Rdot[td + (int64_t)min(wave + NWF * u, ncl - 1) * DSTRIDE];
  } else {
#pragma unroll
    for (int u = 0; u < PB; ++u) pd[u] = 0.0;
  }
#pragma unroll
  for (int u = 0; u < PB; ++u) ph[u] = a.Rpart[rc + (int64_t)min(wave + NWF * u, ncl - 1) * a.lda];
  const int nrn = noright ? 0 : a.nrn;
  constexpr int NRN = 4;  // partial norms per lane in the first batch (256)
  double c = 0.0, ut = 0.0, rk1 = 0.0, rw = 0.0, w1 = 0.0, yk = 0.0, vk = 0.0;
  double rn[NRN];
  const int jw = k + 1 + blockIdx.x * FIN + lane;  // this lane's column of the reflector
  if (wave == 0) {
    c = next ? (double)a.A[rc + (int64_t)(k + 1) * a.lda] : 0.0;
    ut = (double)Ur[(int64_t)t * a.lda];
    rk1 = a.rvec[k + 1];
    rw = a.rvec[jw < a.n ? jw : k + 1];
    w1 = (td <= t) ? a.Y[(k + 1) + (int64_t)td * a.ldy] : a.V[(k + 1) + (int64_t)(td - t - 1) * a.ldy];
#pragma unroll
    for (int u = 0; u < NRN; ++u)  // used if lane + 64 u < nrn
      rn[u] = a.rnorm[min(lane + FIN * u, max(nrn - 1, 0))];
  } else if (wave == 1) {
    const int tk = min(lane, t);
    yk = a.Y[(k + 1) + (int64_t)tk * a.ldy];
    vk = a.V[(k + 1) + (int64_t)tk * a.ldy];
  }
  __builtin_amdgcn_sched_barrier(0);

  // partial sums of this wave's chunks, in chunk order
  double h = 0.0, ds = 0.0;
  if (!noright) {
#pragma unroll
    for (int u = 0; u < PB; ++u)
      if (wave + NWF * u < nccR) {
        ds += pd[u];
        h += ph[u];
      }
    for (int q0 = NWF * PB; q0 < nccR; q0 += NWF * PB) {
#pragma unroll
      for (int u = 0; u < PB; ++u) {
        const int q = min(q0 + wave + NWF * u, nccR - 1);
        pd[u] = a.Rdot[td + (int64_t)q * DSTRIDE];
        ph[u] = a.Rpart[rc + (int64_t)q * a.lda];
      }
#pragma unroll
      for (int u = 0; u < PB; ++u)
        if (q0 + wave + NWF * u < nccR) {
          ds += pd[u];
          h += ph[u];
        }
    }
  }
  hp[wave][lane] = h;
  dp[wave][lane] = ds;
  Refl R;
  R.nrm = R.sgn = R.rl = 0.0;
  if (wave == 0) {
    double r2 = 0.0;
#pragma unroll
    for (int u = 0; u < NRN; ++u)
      if (lane + FIN * u < nrn) r2 += rn[u];
    for (int i = lane + FIN * NRN; i < nrn; i += FIN) r2 += a.rnorm[i];
    if (!noright) R.make(wave_sum(r2), rk1);
    if (lane == 0) {
      rfl[0] = R.nrm;
      rfl[1] = R.sgn;
      rfl[2] = R.rl;
      vk1[t] = noright ? 0.0 : R.vec(rk1, true);
    }
  } else if (wave == 1) {
    if (lane <= t) yk1[lane] = yk;
    if (lane < t) vk1[lane] = vk;
  }
  __syncthreads();
  if (wave == 0 && dotr) {
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < NWF; ++q) d += dp[q][lane];
    dyv[lane] = R.dot(d, w1);
  }
  __syncthreads();

  // this wave's panel terms of x and c
  double xc = 0.0, cc = 0.0;
#pragma unroll
  for (int i = 0; i < PW; ++i) {
    const int s = s0 + i;
    if (s < t) {
      if (!noright) xc += us[i] * dyv[s] + xs[i] * dyv[t + 1 + s];
      cc += us[i] * yk1[s] + xs[i] * vk1[s];
    }
  }
  xp[wave][lane] = xc;
  cp[wave][lane] = cc;
  __syncthreads();
  double xrow = 0.0;  // dots: x_t on this lane's row (wave 0)
  if (wave == 0) {
  double cc2 = 0.0, cnew = 0.0;
  if (rv) {
    double hs = 0.0, xs2 = 0.0, cs2 = 0.0;
#pragma unroll
    for (int q = 0; q < NWF; ++q) {
      hs += hp[q][lane];
      xs2 += xp[q][lane];
      cs2 += cp[q][lane];
    }
    double x = R.dot(hs, c);  // 0 if noright
    x -= xs2;
    if (!noright) {
      x -= ut * dyv[t];
      x *= 2.0;
    }
    a.X[row + (int64_t)t * a.lda] = x;
    xrow = x;
    if (next) {
      c -= cs2;
      c -= ut * yk1[t] + x * vk1[t];
      a.cvec[row] = c;
      cc2 = c * c;
      cnew = c;
    }
  }
  if (dots) cnl[lane] = cnew;
  const double s = wave_sum(cc2);
  if (lane == 0) a.cnorm[blockIdx.x] = s;
  // the right reflector to row k of A and to column t of V (no work-group
  // of this launch reads either: rows > k of A, V[k+1, s < t])
  if (!noright) {
    if (jw < a.n) {
      const double v = R.vec(rw, jw == k + 1);
      a.A[k + (int64_t)jw * a.lda] = (TS)v;
      a.V[jw + (int64_t)t * a.ldy] = v;
    }
    if (blockIdx.x == 0 && lane == 0) a.beta[k] = R.diag();
  }
  if (noright && blockIdx.x == 0 && lane == 0) {
    a.beta[k] = a.rvec[k + 1];
    a.A[k + (int64_t)(k + 1) * a.lda] = (TS)0;
    a.V[(k + 1) + (int64_t)t * a.ldy] = 0.0;
  }
  }
  if (!dots) return;
  __syncthreads();
  // the next step's panel dots over this group's rows (wave w: s = 8w ..
  // 8w+7 below t from us, xs; wave 0 also s = t from u_t and x_t)
  const double cl = cnl[lane];
  double gd[2 * PW];
#pragma unroll
  for (int i = 0; i < PW; ++i) {
    const bool in = s0 + i < t;
    gd[i] = in ? us[i] * cl : 0.0;
    gd[PW + i] = in ? xs[i] * cl : 0.0;
  }
  const double tot = reduce_scatter<double, 2 * PW>(gd);  // lane l: slot l / (64 / 2 PW)
  double* Ld = a.Ldot + (int64_t)blockIdx.x * (2 * NB);
  constexpr int LPS = 64 / (2 * PW);
  if ((lane & (LPS - 1)) == 0) {
    const int slot = lane / LPS, i = slot % PW, sx = s0 + i;
    if (sx < t) Ld[slot < PW ? sx : t + 1 + sx] = tot;
  }
  if (wave == 0) {
    const double du = wave_sum(ut * cl), dx = wave_sum(xrow * cl);
    if (lane == 0) {
      Ld[t] = du;
      Ld[2 * t + 1] = dx;
    }
  }
}

// Last column (m >= n): left reflector only, the 1 x 1 case included
// (bidiag_par.c:384-397).
template <typename TS>
__global__ __launch_bounds__(256) void bd_final_left(StepArgs<TS> a) {
  __shared__ double s_scal;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = a.k;
  if (wave == 0) {
    double s = sum_partials(a.cnorm, a.ncn);
    if (lane == 0) s_scal = s;
  }
  __syncthreads();
  Refl R;
  R.make(s_scal, a.cvec[k]);
  const int row = k + blockIdx.x * 256 + tid;
  if (row < a.m) a.A[row + (int64_t)k * a.lda] = (TS)R.vec(a.cvec[row], row == k);
  if (blockIdx.x == 0 && tid == 0) a.alpha[k] = R.diag();
}

// Wide matrices (m < n), the last row k = m-1 (bidiag.c:133-160,
// bidiag_par.c:361-383): no left reflection -- alpha[k] is M[k, k], which is
// then set to zero -- and the right reflector of row k over columns
// k+1..n-1, giving the N x (N+1) upper bidiagonal.  Phase 1: row k of M on
// those columns (A0[k, j] minus the panel terms of the earlier steps of the
// panel) and a partial |r|^2 per work-group.
template <typename TS>
__global__ __launch_bounds__(FIN) void bd_wide_row(StepArgs<TS> a) {
  __shared__ double uk[NB], xk[NB];
  const int tid = threadIdx.x;
  const int k = a.k, t = a.t;
  const int j = k + 1 + blockIdx.x * FIN + tid;
  const bool jv = j < a.n;
  const int jc = jv ? j : k + 1;  // k + 1 <= n - 1: a valid column
  if (tid < t) {
    uk[tid] = (double)a.A[k + (int64_t)(a.p + tid) * a.lda];
    xk[tid] = a.X[k + (int64_t)tid * a.lda];
  }
  __syncthreads();
  double r = (double)a.A[k + (int64_t)jc * a.lda];
  for (int s = 0; s < t; ++s)
    r -= uk[s] * a.Y[jc + (int64_t)s * a.ldy] + xk[s] * a.V[jc + (int64_t)s * a.ldy];
  if (jv) a.rvec[j] = r;
  const double s2 = wave_sum(jv ? r * r : 0.0);
  if (tid == 0) a.rnorm[blockIdx.x] = s2;
}

// Phase 2: the reflector into A[k, k+1:n], beta[k], alpha[k] = M[k, k]
// (cvec[k], the column the previous step's right_fin left) and A[k, k] = 0.
template <typename TS>
__global__ __launch_bounds__(FIN) void bd_wide_fin(StepArgs<TS> a, int nrn) {
  const int tid = threadIdx.x;
  const int k = a.k;
  double rn = 0.0;
  for (int i = tid; i < nrn; i += FIN) rn += a.rnorm[i];
  Refl R;
  R.make(wave_sum(rn), a.rvec[k + 1]);
  const int j = k + 1 + blockIdx.x * FIN + tid;
  if (j < a.n) a.A[k + (int64_t)j * a.lda] = (TS)R.vec(a.rvec[j], j == k + 1);
  if (blockIdx.x == 0 && tid == 0) {
    a.beta[k] = R.diag();
    a.alpha[k] = a.cvec[k];
    a.A[k + (int64_t)k * a.lda] = (TS)0;
  }
}

// ---------------------------------------------------------------------------

struct BidiagBufs {
  int ldy, nchL_max, nccR_max, ncn_max, nrn_max;
};

static BidiagBufs bidiag_dims(int m, int n, int lda) {
  BidiagBufs d;
  d.ldy = (int)round_up(n, 16);
  // left GEMV partials: one per 2048-row chunk, or one per 128-row tile of
  // the trailing update that carries the next column's left GEMV
  d.nchL_max = (int)std::max({ceil_div(lda, RC), ceil_div(lda, TRAIL_TILE),
                              ceil_div(std::max(lda, n), FIN)}) + 1;  // also right_fin's dots
  d.nccR_max = (int)std::max<int64_t>(ceil_div(n, 64), FG_GMAX) + 1;  // chunks or fused groups
  d.ncn_max = (int)std::max(ceil_div(m, 256), ceil_div(std::max(m, n), FIN)) + 1;
  d.nrn_max = (int)std::max<int64_t>(ceil_div(n, FIN), FG_GMAX) + 1;
  return d;
}

// ---- operands of the reduced-precision trailing update ----------------------

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Entry (i, c) of the update operand, i < rows, c < 2 NB:
//   side 0: [U X] (rows r0.. of the panel's reflectors, stored as float, and X)
//   side 1: [Y V] (contiguous double columns)
__device__ __forceinline__ float upd_entry(int side, int i, int c, const float* U, int ldu,
                                           const double* D, int ldd) {
  if (side == 0 && c < NB) return U[i + (int64_t)c * ldu];
  const int cc = side == 0 ? c - NB : c;
  return (float)D[i + (int64_t)cc * ldd];
}

// FP32 copy of an update operand (rows x 2NB, ld ldo).
__global__ void upd_pack_f32(int side, int rows, const float* U, int ldu, const double* D,
                             int ldd, float* out, int ldo) {
  const int64_t total = (int64_t)rows * 2 * NB;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / rows), i = (int)(e - (int64_t)c * rows);
    out[i + (int64_t)c * ldo] = upd_entry(side, i, c, U, ldu, D, ldd);
  }
}

// BF16 operand.  pieces = 1: hi parts (K = 2NB).  pieces = 3 (split BF16):
// K = 6NB laid out so that [hi hi lo] (side 0) times [hi lo hi]^T (side 1)
// sums hi*hi + hi*lo + lo*hi, x = hi + lo the two-term BF16 split.
__global__ void upd_pack_bf16(int side, int pieces, int rows, const float* U, int ldu,
                              const double* D, int ldd, unsigned short* out, int ldo) {
  const int64_t total = (int64_t)rows * 2 * NB;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / rows), i = (int)(e - (int64_t)c * rows);
    const float x = upd_entry(side, i, c, U, ldu, D, ldd);
    const unsigned short hi = bf16_rne(x);
    out[i + (int64_t)c * ldo] = hi;
    if (pieces == 3) {
      const unsigned short lo = bf16_rne(x - bf16_val(hi));
      out[i + (int64_t)(c + 2 * NB) * ldo] = side == 0 ? hi : lo;
      out[i + (int64_t)(c + 4 * NB) * ldo] = side == 0 ? lo : hi;
    }
  }
}

// ---------------------------------------------------------------------------

static void bidiag_plan_common(ArenaPlan& p, int m, int n, int lda) {
  BidiagBufs d = bidiag_dims(m, n, lda);
  p.add<double>((size_t)d.ldy * NB * 2);   // Y, V
  p.add<double>((size_t)lda * NB * 2);     // copy of the panel's U, then X
  p.add<double>(lda);                      // cvec
  p.add<double>(d.ldy);                    // rvec
  p.add<double>((size_t)d.nchL_max * d.ldy);
  p.add<double>((size_t)d.nccR_max * lda);
  p.add<double>((size_t)d.nchL_max * 2 * NB);
  p.add<double>((size_t)d.nccR_max * DSTRIDE);
  p.add<double>(d.ncn_max);
  p.add<double>(d.nrn_max);
}

void bidiag_plan(ArenaPlan& p, int m, int n, int lda) { bidiag_plan_common(p, m, n, lda); }

void bidiag_plan_f32(ArenaPlan& p, int m, int n, int lda, int upd) {
  bidiag_plan_common(p, m, n, lda);
  const int ldy = (int)round_up(n, 16);
  if (upd == UPD_F32) {
    p.add<float>((size_t)lda * 2 * NB);
    p.add<float>((size_t)ldy * 2 * NB);
  } else {
    const int pieces = upd == UPD_BF16X3 ? 3 : 1;
    p.add<unsigned short>((size_t)lda * 2 * NB * pieces);
    p.add<unsigned short>((size_t)ldy * 2 * NB * pieces);
  }
}

// Trailing widths (columns right of k) up to which the right GEMV takes
// 64-column chunks: there 128-column chunks leave too few work-groups to
// keep HBM busy.  Per 8192^2 SVD, by threshold: 0 (never) 654.3 ms, 1024
// 652.2, 2048 650.2, 3072 650.8, 4096 649.7, 8192 (always) 656.9, where
// the right GEMV itself is fastest (30.8 against 32.3 us per launch) but
// bd_right_fin sums 128 chunk partials per row at the largest widths.
constexpr int RIGHT_NARROW_COLS = 4096;

// The right GEMV of step a.k; returns its chunk count.  16 columns are
// prefetched before the prologue (8: 1.5-2 ms more per 8192^2 SVD; 24
// would lower the occupancy to two waves per SIMD).
template <typename TS>
static int launch_right_gemv(const StepArgs<TS>& a, int nchL, bool clock, hipStream_t st) {
  const int k = a.k;
  const int rbr = ((k + 1) / 64) * 64;
  const int nrb = (int)ceil_div(a.m - rbr, RBR);
  const bool narrow = a.n - k - 1 <= RIGHT_NARROW_COLS;
  const int ncc = (int)ceil_div(a.n - k - 1, narrow ? 64 : CCR);
  auto kern = narrow ? bd_right_gemv<TS, 64, 16> : bd_right_gemv<TS, CCR, 16>;
  hipEvent_t ke0, ke1;
  if (clock && kclock_on() && kclock_pair(KCLOCK_RIGHT_GEMV, &ke0, &ke1) == 0)
    hipExtLaunchKernelGGL(kern, dim3(nrb + 1, ncc), dim3(256), 0, st, ke0, ke1, 0, a, nrb, rbr,
                          nchL);
  else
    hipLaunchKernelGGL(kern, dim3(nrb + 1, ncc), dim3(256), 0, st, a, nrb, rbr, nchL);
  return ncc;
}

static bool fused_gemv_on() {
  static const bool on = getenv("DDCSVD_TWO_PASS") == nullptr;
  return on;
}

template <typename TS>
static int bidiag_drive(int m, int n, TS* A, int lda, double* alpha, double* beta, int upd,
                        Arena& ws, hipStream_t st) {
  if (m < 1 || n < 1 || lda < m || (lda % 2) != 0) return DDC_E_ARG;
  BidiagBufs d = bidiag_dims(m, n, lda);
  StepArgs<TS> a;
  a.m = m;
  a.n = n;
  a.lda = lda;
  a.ldy = d.ldy;
  a.A = A;
  a.alpha = alpha;
  a.beta = beta;
  a.Y = ws.take<double>((size_t)d.ldy * NB * 2);
  a.V = a.Y + (size_t)d.ldy * NB;
  // [U X] and [Y V]: the trailing update's two dense operands
  double* UX = ws.take<double>((size_t)lda * NB * 2);
  a.X = UX ? UX + (size_t)lda * NB : nullptr;
  // FP64 storage: left_gemv also writes each reflector next to X, so that
  // [U X] is one dense operand of the trailing update
  a.Ucopy = sizeof(TS) == 8 ? UX : nullptr;
  a.cvec = ws.take<double>(lda);
  a.rvec = ws.take<double>(d.ldy);
  a.Lpart = ws.take<double>((size_t)d.nchL_max * d.ldy);
  a.Rpart = ws.take<double>((size_t)d.nccR_max * lda);
  a.Ldot = ws.take<double>((size_t)d.nchL_max * 2 * NB);
  a.Rdot = ws.take<double>((size_t)d.nccR_max * DSTRIDE);
  a.cnorm = ws.take<double>(d.ncn_max);
  a.rnorm = ws.take<double>(d.nrn_max);
  if (!a.rnorm) return DDC_E_NOMEM;
  float *opA32 = nullptr, *opB32 = nullptr;
  unsigned short *opA16 = nullptr, *opB16 = nullptr;
  const int pieces = upd == UPD_BF16X3 ? 3 : 1;
  if constexpr (sizeof(TS) == 4) {
    if (upd == UPD_F32) {
      opA32 = ws.take<float>((size_t)lda * 2 * NB);
      opB32 = ws.take<float>((size_t)d.ldy * 2 * NB);
      if (!opB32) return DDC_E_NOMEM;
    } else {
      opA16 = ws.take<unsigned short>((size_t)lda * 2 * NB * pieces);
      opB16 = ws.take<unsigned short>((size_t)d.ldy * 2 * NB * pieces);
      if (!opB16) return DDC_E_NOMEM;
    }
  }

  DDC_HIP(hipMemsetAsync(a.Y, 0, sizeof(double) * (size_t)d.ldy * NB * 2, st));
  DDC_HIP(hipMemsetAsync(a.X, 0, sizeof(double) * (size_t)lda * NB, st));
  DDC_HIP(hipMemsetAsync(beta, 0, sizeof(double) * n, st));

  a.ncn = (int)ceil_div(m, 256);
  a.nrn = 0;
  hipLaunchKernelGGL(bd_init<TS>, dim3(a.ncn), dim3(256), 0, st, m, A, a.cvec, a.cnorm);
  DDC_LAUNCHED();

  const int mn = std::min(m, n);
  // column steps whose two GEMVs are one pass (bd_fused_gemv): the plain
  // steps (not the last two columns or the wide last row) on at most
  // FG_ROWS rows
  auto one_pass = [&](int kk) {
    return fused_gemv_on() && kk < n - 2 && !(m < n && kk >= m - 1) && m - kk <= FG_ROWS &&
           n - kk - 1 <= FG_GMAX * FG_MAXCOL;
  };
  int fusedL = 0;  // tiles of left GEMV partials the trailing update left for step k, or 0
  int nfr_prev = 0;  // groups of the previous bd_right_fin (rows of its panel dots)
  for (int k = 0; k < mn; ++k) {
    a.k = k;
    a.p = (k / NB) * NB;
    a.t = k - a.p;
    const int nt = fusedL;
    fusedL = 0;
    if (m < n && k == m - 1) {
      const int nw = (int)ceil_div(n - k - 1, FIN);
      hipLaunchKernelGGL(bd_wide_row<TS>, dim3(nw), dim3(FIN), 0, st, a);
      DDC_LAUNCHED();
      hipLaunchKernelGGL(bd_wide_fin<TS>, dim3(nw), dim3(FIN), 0, st, a, nw);
      DDC_LAUNCHED();
      break;
    }
    if (k == n - 1) {
      hipLaunchKernelGGL(bd_final_left<TS>, dim3((unsigned)ceil_div(m - k, 256)), dim3(256), 0,
                         st, a);
      DDC_LAUNCHED();
      break;
    }
    // right_fin's grid covers the rows k+1..m-1 and, for the right
    // reflector, the columns k+1..n-1
    const int nfr = (int)ceil_div(std::max(m, n) - k - 1, FIN);
    // the next step of this panel, if one pass, takes its panel dots from
    // this step's bd_right_fin
    const int dots = (a.t + 1 < NB && one_pass(k + 1)) ? 1 : 0;
    if (one_pass(k)) {
      const int G = std::min(FG_GMAX, n - k - 1);
      launch_fused_gemv(a, G, a.t > 0 ? nfr_prev : 0, true, st);
      DDC_LAUNCHED();
      a.nrn = G;
      // the one-pass groups' partials: up to 256, one batch of 32 per wave of 8
      if (G > 4 * 16)
        hipLaunchKernelGGL((bd_right_fin<TS, 32, 8>), dim3(nfr), dim3(8 * FIN), 0, st, a, G, 0,
                           dots);
      else
        hipLaunchKernelGGL((bd_right_fin<TS, 16, 4>), dim3(nfr), dim3(4 * FIN), 0, st, a, G, 0,
                           dots);
      DDC_LAUNCHED();
    } else {
    // left reflector of column k and the left GEMV
    const int rb0 = (k / 64) * 64;
    const int ncb = (int)ceil_div(n - k - 1, CBL);
    const int ndl = (int)ceil_div(2 * a.t, DPW);
    int nchL = (int)ceil_div(m - rb0, RC);
    hipEvent_t ke0, ke1;
    if (nt > 0) {
      // the trailing update formed the partial sums (t = 0: no panel dots)
      const int ncs = (int)ceil_div(n - k - 1, 256);
      hipLaunchKernelGGL(bd_left_red<TS>, dim3(ncs + (unsigned)ceil_div(m - k, 256)), dim3(256),
                         0, st, a, nt, ncs);
      nchL = 1;
    } else if (kclock_on() && kclock_pair(KCLOCK_LEFT_GEMV, &ke0, &ke1) == 0)
      hipExtLaunchKernelGGL((bd_left_gemv<TS, RC / 512>), dim3(ncb + ndl, nchL), dim3(256), 0,
                            st, ke0, ke1, 0, a, ncb, rb0);
    else
      hipLaunchKernelGGL((bd_left_gemv<TS, RC / 512>), dim3(ncb + ndl, nchL), dim3(256), 0, st,
                         a, ncb, rb0);
    DDC_LAUNCHED();
    // row k of M (the left finalization), the right reflector of row k and
    // the right GEMV
    if (k == n - 2) {
      const int nfl = (int)ceil_div(n - k - 1, FIN);
      hipLaunchKernelGGL(bd_left_fin<TS>, dim3(nfl), dim3(FIN), 0, st, a, nchL);
      DDC_LAUNCHED();
      a.nrn = nfl;
      hipLaunchKernelGGL((bd_right_fin<TS, 16, 4>), dim3(nfr), dim3(4 * FIN), 0, st, a, 0, 1, 0);
      DDC_LAUNCHED();
    } else {
      const int ncc = launch_right_gemv(a, nchL, true, st);
      DDC_LAUNCHED();
      a.nrn = ncc;
      hipLaunchKernelGGL((bd_right_fin<TS, 16, 4>), dim3(nfr), dim3(4 * FIN), 0, st, a, ncc, 0,
                         dots);
      DDC_LAUNCHED();
    }
    }
    a.ncn = nfr;
    nfr_prev = nfr;
    // blocked trailing update at the end of a full panel
    if (a.t == NB - 1 && a.p + NB < n) {
      const int r0 = a.p + NB;
      TS* C = A + r0 + (int64_t)r0 * lda;
      if constexpr (sizeof(TS) == 8) {
        // The update also takes the partial sums of the next step's left
        // GEMV (column r0 of the updated matrix is c, from bd_right_fin)
        // when that step is a plain one: one pass over the trailing matrix
        // less per panel.
        const bool fuse = !one_pass(r0) && r0 < n - 1 && !(m < n && r0 == m - 1);
        DDC_TRY(dgemm_trail64(m - r0, n - r0, UX + r0, lda, a.Y + r0, d.ldy, C, lda,
                              fuse ? a.cvec + r0 : nullptr, a.Lpart + r0, d.ldy, 1, st));
        if (fuse) fusedL = (int)ceil_div(m - r0, TRAIL_TILE);
      } else {
        const float* Up = A + r0 + (int64_t)a.p * lda;
        const int ra = m - r0, rb = n - r0;
        const unsigned ga = (unsigned)std::min<int64_t>(ceil_div((int64_t)ra * 2 * NB, 256), 4096);
        const unsigned gb = (unsigned)std::min<int64_t>(ceil_div((int64_t)rb * 2 * NB, 256), 4096);
        if (upd == UPD_F32) {
          hipLaunchKernelGGL(upd_pack_f32, dim3(ga), dim3(256), 0, st, 0, ra, Up, lda,
                             a.X + r0, lda, opA32, lda);
          hipLaunchKernelGGL(upd_pack_f32, dim3(gb), dim3(256), 0, st, 1, rb, Up, lda,
                             a.Y + r0, d.ldy, opB32, d.ldy);
          DDC_LAUNCHED();
          DDC_TRY(sgemm('N', 'T', ra, rb, 2 * NB, -1.0f, opA32, lda, opB32, d.ldy, 1.0f, C, lda,
                        st));
        } else {
          hipLaunchKernelGGL(upd_pack_bf16, dim3(ga), dim3(256), 0, st, 0, pieces, ra, Up, lda,
                             a.X + r0, lda, opA16, lda);
          hipLaunchKernelGGL(upd_pack_bf16, dim3(gb), dim3(256), 0, st, 1, pieces, rb, Up, lda,
                             a.Y + r0, d.ldy, opB16, d.ldy);
          DDC_LAUNCHED();
          DDC_TRY(gemm_bf16_f32('N', 'T', ra, rb, 2 * NB * pieces, -1.0f, opA16, lda, opB16,
                                d.ldy, 1.0f, C, lda, st));
        }
      }
    }
  }
  return DDC_OK;
}

int bidiag_run(int m, int n, double* A, int lda, double* alpha, double* beta, Arena& ws,
               hipStream_t st) {
  return bidiag_drive<double>(m, n, A, lda, alpha, beta, 0, ws, st);
}

// Measurement: `steps` left-GEMV launches, at column steps k = 0..steps-1,
// on the panel-start matrix A (m >= n; A is overwritten with reflectors,
// alpha with their diagonals).  The column of step 0 is c for every step and
// no other kernel of the step runs, so a counter pass sees the left GEMV's
// own traffic for the trailing blocks A[k:m, k+1:n].
// Measurement: `steps` one-pass GEMV launches, at column steps k =
// 0..steps-1, on the panel-start matrix A (m >= n, m <= FG_ROWS; A is
// overwritten with reflectors, alpha with their diagonals).  The column of
// step 0 is c for every step, the panel dots are zero and no other kernel
// of the step runs, so a counter pass sees the kernel's own traffic for
// the trailing blocks A[k:m, k+1:n] (one read each).
int bidiag_fused_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps,
                           Arena& ws, hipStream_t st) {
  if (m < n || n < 3 || m > FG_ROWS || lda < m || (lda % 2) != 0 || steps < 1 || steps > n - 2)
    return DDC_E_ARG;
  BidiagBufs d = bidiag_dims(m, n, lda);
  StepArgs<double> a;
  a.m = m;
  a.n = n;
  a.lda = lda;
  a.ldy = d.ldy;
  a.A = A;
  a.alpha = alpha;
  a.beta = nullptr;
  a.Y = ws.take<double>((size_t)d.ldy * NB * 2);
  a.V = a.Y + (size_t)d.ldy * NB;
  double* UX = ws.take<double>((size_t)lda * NB * 2);
  a.X = UX ? UX + (size_t)lda * NB : nullptr;
  a.Ucopy = UX;
  a.cvec = ws.take<double>(lda);
  a.rvec = ws.take<double>(d.ldy);
  a.Lpart = ws.take<double>((size_t)d.nchL_max * d.ldy);
  a.Rpart = ws.take<double>((size_t)d.nccR_max * lda);
  a.Ldot = ws.take<double>((size_t)d.nchL_max * 2 * NB);
  a.Rdot = ws.take<double>((size_t)d.nccR_max * DSTRIDE);
  a.cnorm = ws.take<double>(d.ncn_max);
  a.rnorm = ws.take<double>(d.nrn_max);
  if (!a.rnorm) return DDC_E_NOMEM;
  DDC_HIP(hipMemsetAsync(a.Y, 0, sizeof(double) * (size_t)d.ldy * NB * 2, st));
  DDC_HIP(hipMemsetAsync(a.X, 0, sizeof(double) * (size_t)lda * NB, st));
  a.ncn = (int)ceil_div(m, 256);
  a.nrn = 0;
  hipLaunchKernelGGL(bd_init<double>, dim3(a.ncn), dim3(256), 0, st, m, A, a.cvec, a.cnorm);
  DDC_LAUNCHED();
  for (int k = 0; k < steps; ++k) {
    a.k = k;
    a.p = (k / NB) * NB;
    a.t = k - a.p;
    launch_fused_gemv(a, std::min(FG_GMAX, n - k - 1), 0, false, st);
    DDC_LAUNCHED();
  }
  return DDC_OK;
}

int bidiag_left_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps, Arena& ws,
                          hipStream_t st) {
  if (m < n || n < 2 || lda < m || (lda % 2) != 0 || steps < 1 || steps > n - 1)
    return DDC_E_ARG;
  BidiagBufs d = bidiag_dims(m, n, lda);
  StepArgs<double> a;
  a.m = m;
  a.n = n;
  a.lda = lda;
  a.ldy = d.ldy;
  a.A = A;
  a.alpha = alpha;
  a.beta = nullptr;
  a.Y = ws.take<double>((size_t)d.ldy * NB * 2);
  a.V = a.Y + (size_t)d.ldy * NB;
  double* UX = ws.take<double>((size_t)lda * NB * 2);
  a.X = UX ? UX + (size_t)lda * NB : nullptr;
  a.Ucopy = UX;
  a.cvec = ws.take<double>(lda);
  a.rvec = ws.take<double>(d.ldy);
  a.Lpart = ws.take<double>((size_t)d.nchL_max * d.ldy);
  a.Rpart = ws.take<double>((size_t)d.nccR_max * lda);
  a.Ldot = ws.take<double>((size_t)d.nchL_max * 2 * NB);
  a.Rdot = ws.take<double>((size_t)d.nccR_max * DSTRIDE);
  a.cnorm = ws.take<double>(d.ncn_max);
  a.rnorm = ws.take<double>(d.nrn_max);
  if (!a.rnorm) return DDC_E_NOMEM;
  DDC_HIP(hipMemsetAsync(a.X, 0, sizeof(double) * (size_t)lda * NB, st));
  a.ncn = (int)ceil_div(m, 256);
  a.nrn = 0;
  hipLaunchKernelGGL(bd_init<double>, dim3(a.ncn), dim3(256), 0, st, m, A, a.cvec, a.cnorm);
  DDC_LAUNCHED();
  for (int k = 0; k < steps; ++k) {
    a.k = k;
    a.p = (k / NB) * NB;
    a.t = k - a.p;
    const int rb0 = (k / 64) * 64;
    const int ncb = (int)ceil_div(n - k - 1, CBL);
    const int ndl = (int)ceil_div(2 * a.t, DPW);
    const int nchL = (int)ceil_div(m - rb0, RC);
    hipLaunchKernelGGL((bd_left_gemv<double, RC / 512>), dim3(ncb + ndl, nchL), dim3(256), 0, st,
                       a, ncb, rb0);
    DDC_LAUNCHED();
  }
  return DDC_OK;
}

// Measurement: `steps` right-GEMV launches, at column steps k = 0..steps-1,
// after one left GEMV at k = 0 (m >= n; A's column 0 is overwritten with a
// reflector).  Every launch reads the partial sums of that left GEMV and
// the zeroed panel buffers, so no other kernel runs between them and a
// counter pass sees the right GEMV's own traffic for the trailing blocks
// A[k+1:m, k+1:n]; the values computed are discarded.
int bidiag_right_gemv_only(int m, int n, double* A, int lda, double* alpha, int steps,
                           Arena& ws, hipStream_t st) {
  if (m < n || n < 3 || lda < m || (lda % 2) != 0 || steps < 1 || steps > n - 2)
    return DDC_E_ARG;
  BidiagBufs d = bidiag_dims(m, n, lda);
  StepArgs<double> a;
  a.m = m;
  a.n = n;
  a.lda = lda;
  a.ldy = d.ldy;
  a.A = A;
  a.alpha = alpha;
  a.beta = nullptr;
  a.Y = ws.take<double>((size_t)d.ldy * NB * 2);
  a.V = a.Y + (size_t)d.ldy * NB;
  double* UX = ws.take<double>((size_t)lda * NB * 2);
  a.X = UX ? UX + (size_t)lda * NB : nullptr;
  a.Ucopy = UX;
  a.cvec = ws.take<double>(lda);
  a.rvec = ws.take<double>(d.ldy);
  a.Lpart = ws.take<double>((size_t)d.nchL_max * d.ldy);
  a.Rpart = ws.take<double>((size_t)d.nccR_max * lda);
  a.Ldot = ws.take<double>((size_t)d.nchL_max * 2 * NB);
  a.Rdot = ws.take<double>((size_t)d.nccR_max * DSTRIDE);
  a.cnorm = ws.take<double>(d.ncn_max);
  a.rnorm = ws.take<double>(d.nrn_max);
  if (!a.rnorm) return DDC_E_NOMEM;
  DDC_HIP(hipMemsetAsync(a.Y, 0, sizeof(double) * (size_t)d.ldy * NB * 2, st));
  DDC_HIP(hipMemsetAsync(UX, 0, sizeof(double) * (size_t)lda * NB * 2, st));
  DDC_HIP(hipMemsetAsync(a.Ldot, 0, sizeof(double) * (size_t)d.nchL_max * 2 * NB, st));
  a.ncn = (int)ceil_div(m, 256);
  a.nrn = 0;
  hipLaunchKernelGGL(bd_init<double>, dim3(a.ncn), dim3(256), 0, st, m, A, a.cvec, a.cnorm);
  DDC_LAUNCHED();
  a.k = a.p = a.t = 0;
  const int ncb0 = (int)ceil_div(n - 1, CBL);
  const int nchL0 = (int)ceil_div(m, RC);
  hipLaunchKernelGGL((bd_left_gemv<double, RC / 512>), dim3(ncb0, nchL0), dim3(256), 0, st, a,
                     ncb0, 0);
  DDC_LAUNCHED();
  for (int k = 0; k < steps; ++k) {
    a.k = k;
    a.p = (k / NB) * NB;
    a.t = k - a.p;
    // the left GEMV's partials cover rows from 0: the chunk count of step k
    // is at most that of step 0
    const int rb0 = (k / 64) * 64;
    const int nchL = (int)ceil_div(m - rb0, RC);
    launch_right_gemv(a, nchL, false, st);
    DDC_LAUNCHED();
  }
  return DDC_OK;
}

int bidiag_run_f32(int m, int n, float* A, int lda, double* alpha, double* beta, int upd,
                   Arena& ws, hipStream_t st) {
  if (upd != UPD_F32 && upd != UPD_BF16X3 && upd != UPD_BF16) return DDC_E_ARG;
  return bidiag_drive<float>(m, n, A, lda, alpha, beta, upd, ws, st);
}

}  // namespace ddc
