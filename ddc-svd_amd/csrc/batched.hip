// batched.hip -- SVD of many small matrices of uneven shapes in one
// persistent launch (FP64 or FP32), replayed from a captured HIP graph.
//
// No reference counterpart: the reference factors one matrix per call
// (svd_gpu.c:53).  Method: one-sided (Hestenes) Jacobi by column blocks,
// one wavefront per matrix.
//
//  * A persistent grid of 4-wave workgroups pulls matrix indices from an
//    atomic counter; the queue is ordered by descending work (n^2 m), so the
//    large matrices start first and the small ones fill the tail.  A matrix
//    with min(m, n) >= coop_nt is factored by a whole workgroup (its waves
//    take disjoint block pairs of each round); from the first smaller matrix
//    on, every wave takes matrices on its own.
//  * A matrix (tall orientation: mt = max(m,n) <= 256 rows, nt = min(m,n)
//    columns; a wide matrix is processed as its transpose) is split into
//    column blocks of width B.  For each block pair (I, J) the wave holds the
//    2B columns of the working matrix and of V in registers (lane l holds rows
//    l, l+64, ...), runs every cross pair (B rounds of B disjoint pairs) and,
//    on the first visit of a block in a sweep, its internal pairs
//    (round-robin, B-1 rounds), then writes the columns back.  Only the dot
//    product gamma needs a wave reduction per rotation: the squared column
//    norms are reduced when a block pair is loaded and then carried through
//    the rotations (alpha' = alpha - t gamma, beta' = beta + t gamma).
//  * Sweeps stop when every pair satisfies |gamma| <= tol sqrt(alpha beta),
//    tol = sqrt(mt) * eps.  Then sigma_j = ||a_j||, U_j = a_j / sigma_j, and
//    the columns are written in ascending order of sigma.  Columns of exactly
//    zero norm are completed to an orthonormal basis by Gram-Schmidt on unit
//    vectors.
//
// Three tunables are picked by measurement (svd_gpu_batched_tune): the kernel
// variant (block width per row class, which also sets the register budget
// and so the occupancy), the number of resident workgroups per CU, and the
// cooperation threshold coop_nt.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "../../include/svd_gpu.h"
#include "ddc_common.h"

namespace ddc {

constexpr int BJ_MAXDIM = 256;  // max(m, n) of one matrix
constexpr int BJ_MAXSWEEP = 40;
constexpr int BJ_WGW = 4;       // waves per workgroup

struct BatchItem {
  int m, n;
  int64_t a, s, u, v;  // element offsets of A, sigma, U, V
  int64_t wa, wv;      // element offsets of the working matrix and of V
};

template <typename T>
struct BjArgs {
  int count;
  const BatchItem* items;
  const int* order;  // processing order
  int* counter;      // queue head; zero before the launch
  const T* A;
  T* S;
  T* U;
  T* V;
  T* W;      // working matrices (tall orientation, scaled)
  T* WV;     // working V
  int* info;  // per matrix: sweeps used, or -1 if not converged (may be null)
  int coop_nt;  // matrices with min(m, n) >= coop_nt are factored by a whole workgroup
};

template <typename T>
struct Num;
template <>
struct Num<double> {
  static constexpr double eps = 1.1102230246251565e-16;
  static constexpr double big = 1e150;
};
template <>
struct Num<float> {
  static constexpr float eps = 5.9604645e-08f;
  static constexpr float big = 1e18f;
};

// Raw buffer access to one matrix: loads outside [0, count) return 0 and
// stores there are dropped by the hardware range check, so masked rows and
// columns need no branches.  Pass idx < 0 for a masked element.
template <typename T>
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const T* p, int count)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), 0, count * (int)sizeof(T),
                                            0x00020000)) {}
  __device__ __forceinline__ static int off(int idx) {
    return idx >= 0 ? idx * (int)sizeof(T) : 0x7fffff00;
  }
  __device__ __forceinline__ T ld(int idx) const;
  __device__ __forceinline__ void st(int idx, T v) const;
};
template <>
__device__ __forceinline__ double Buf<double>::ld(int idx) const {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off(idx), 0, 0));
}
template <>
__device__ __forceinline__ float Buf<float>::ld(int idx) const {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off(idx), 0, 0));
}
template <>
__device__ __forceinline__ void Buf<double>::st(int idx, double v) const {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(
      decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, off(idx), 0, 0);
}
template <>
__device__ __forceinline__ void Buf<float>::st(int idx, float v) const {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off(idx), 0, 0);
}

template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T wmax(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Broadcast of lane l's value (to a wave-uniform register).
__device__ __forceinline__ double rdlane(double x, int l) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float rdlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Jacobi rotation of the pair with squared norms al, be and dot product ga.
// Returns false (no rotation) when the pair is already orthogonal to tol.
template <typename T>
__device__ __forceinline__ bool jrot(T al, T be, T ga, T tol, T& c, T& s, T& t, T& off) {
  const T den = sqrt(al) * sqrt(be);
  if (!(den > T(0)) || !(fabs(ga) > tol * den)) return false;
  off = fmax(off, fabs(ga) / den);
  const T zeta = (be - al) / (T(2) * ga);
  const T az = fabs(zeta);
  const T tm = az > Num<T>::big ? T(0.5) / az : T(1) / (az + sqrt(T(1) + az * az));
  t = zeta < T(0) ? -tm : tm;
  c = T(1) / sqrt(T(1) + t * t);
  s = c * t;
  return true;
}

// The 2B columns of a block pair held by one wave: positions 0..B-1 are block
// I, B..2B-1 block J; lane l holds rows l + 64 e.
template <typename T, int B, int R, int RV>
struct Block {
  T a[2 * B][R];
  T v[2 * B][RV];
};

template <typename T, int B, int R, int RV>
__device__ __forceinline__ void rotate_cols(Block<T, B, R, RV>& k, int p, int q, T c, T s) {
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const T x = k.a[p][e], y = k.a[q][e];
    k.a[p][e] = c * x - s * y;
    k.a[q][e] = s * x + c * y;
  }
#pragma unroll
  for (int e = 0; e < RV; ++e) {
    const T x = k.v[p][e], y = k.v[q][e];
    k.v[p][e] = c * x - s * y;
    k.v[q][e] = s * x + c * y;
  }
}

// Copies column src of block y to column dst of block x (registers).
template <typename T, int B1, int B2, int R, int RV>
__device__ __forceinline__ void move_col(Block<T, B1, R, RV>& x, int dst,
                                         const Block<T, B2, R, RV>& y, int src) {
#pragma unroll
  for (int e = 0; e < R; ++e) x.a[dst][e] = y.a[src][e];
#pragma unroll
  for (int e = 0; e < RV; ++e) x.v[dst][e] = y.v[src][e];
}
template <typename T, int B, int R, int RV>
__device__ __forceinline__ void move_col(Block<T, B, R, RV>& k, int dst, int src) {
  move_col(k, dst, k, src);
}

// Pairs of one round.  Cross rounds (KIND 0) pair position i with B + i;
// internal rounds (KIND 1) pair, inside each half h, position j with B-1-j
// (j < B/2): slot i = h * B/2 + j.  Between rounds the positions rotate so
// that B cross rounds meet every (I, J) pair and B-1 internal rounds every
// pair inside a block (circle method: position 0 stays, 1..B-1 rotate).
template <int B, int KIND>
__host__ __device__ constexpr int pair_p(int i) {
  return KIND == 0 ? i : (i / (B / 2)) * B + i % (B / 2);
}
template <int B, int KIND>
__host__ __device__ constexpr int pair_q(int i) {
  return KIND == 0 ? B + i : (i / (B / 2)) * B + B - 1 - i % (B / 2);
}

// One round: B disjoint rotations.  nrm: lane c holds the squared norm of
// the column at position c.  The dot products are reduce-scattered across
// the wave so that the lanes of group i hold pair i's; each group computes
// its rotation, and the cosines and sines are broadcast.
template <typename T, int B, int R, int RV, int KIND>
__device__ __forceinline__ void jround(Block<T, B, R, RV>& k, T& nrm, bool enI, bool enJ, T tol,
                                       T& off) {
  const int lane = threadIdx.x & 63;
  T g[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const bool en = KIND == 0 || (i < B / 2 ? enI : enJ);
    T d = T(0);
    if (en) {
#pragma unroll
      for (int e = 0; e < R; ++e) d += k.a[pair_p<B, KIND>(i)][e] * k.a[pair_q<B, KIND>(i)][e];
    }
    g[i] = d;
  }
  constexpr int GS = 64 / B;  // lanes per group
  const T gl = reduce_scatter<T, B>(g);
  const int li = lane / GS;  // the pair this lane's group handles
  const int pl = KIND == 0 ? li : (li / (B / 2)) * B + li % (B / 2);
  const int ql = KIND == 0 ? B + li : (li / (B / 2)) * B + B - 1 - li % (B / 2);
  const T al = __shfl(nrm, pl, 64), be = __shfl(nrm, ql, 64);
  T c = T(1), s = T(0), t = T(0);
  const bool act = jrot(al, be, gl, tol, c, s, t, off);
  const T aln = act ? fmax(al - t * gl, T(0)) : al;
  const T ben = act ? be + t * gl : be;
  // new norms back to the column lanes
  int slot;
  bool qside;
  if (KIND == 0) {
    qside = lane >= B;
    slot = qside ? lane - B : lane;
  } else {
    const int h = lane / B, x = lane % B;
    qside = x >= B / 2;
    slot = h * (B / 2) + (qside ? B - 1 - x : x);
  }
  slot = lane < 2 * B ? slot : 0;
  const T np = __shfl(aln, slot * GS, 64), nq = __shfl(ben, slot * GS, 64);
  if (lane < 2 * B) nrm = qside ? nq : np;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const T si = rdlane(s, i * GS);
    if (si != T(0))
      rotate_cols(k, pair_p<B, KIND>(i), pair_q<B, KIND>(i), rdlane(c, i * GS), si);
  }
}

// Position rotations between rounds (registers and norms).
template <typename T, int B, int R, int RV>
__device__ __forceinline__ void shift_cross(Block<T, B, R, RV>& k, T& nrm) {
  const int lane = threadIdx.x & 63;
  Block<T, 1, R, RV> tmp;
  move_col(tmp, 0, k, B);
#pragma unroll
  for (int x = 0; x < B - 1; ++x) move_col(k, B + x, B + x + 1);
  move_col(k, 2 * B - 1, tmp, 0);
  const int src = (lane >= B && lane < 2 * B) ? B + (lane - B + 1) % B : lane;
  nrm = __shfl(nrm, src, 64);
}

template <typename T, int B, int R, int RV>
__device__ __forceinline__ void shift_intra(Block<T, B, R, RV>& k, T& nrm) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int h = 0; h < 2 * B; h += B) {
    Block<T, 1, R, RV> tmp;
    move_col(tmp, 0, k, h + 1);
#pragma unroll
    for (int x = 1; x < B - 1; ++x) move_col(k, h + x, h + x + 1);
    move_col(k, h + B - 1, tmp, 0);
  }
  const int x = lane % B;
  const int src = (lane < 2 * B && x >= 1) ? lane - x + 1 + x % (B - 1) : lane;
  nrm = __shfl(nrm, src, 64);
}

// One block-pair step.  cI / cJ: first columns of the blocks (cJ < 0: block I
// alone).  Returns the largest relative dot product seen.
template <typename T, int B, int R, int RV>
__device__ __forceinline__ T block_step(T* __restrict__ W, T* __restrict__ WV, int mt, int nt, int cI, int cJ,
                        bool intraI, bool intraJ, T tol) {
  const int lane = threadIdx.x & 63;
  const Buf<T> bw(W, mt * nt), bv(WV, nt * nt);
  Block<T, B, R, RV> k;
  T part[2 * B];
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) {
    const int col = c < B ? cI + c : cJ + (c - B);
    const bool ok = (c < B || cJ >= 0) && col < nt;
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int r = lane + 64 * e;
      k.a[c][e] = bw.ld(ok && r < mt ? r + col * mt : -1);
    }
#pragma unroll
    for (int e = 0; e < RV; ++e) {
      const int r = lane + 64 * e;
      k.v[c][e] = bv.ld(ok && r < nt ? r + col * nt : -1);
    }
  }
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) {
    T s = T(0);
#pragma unroll
    for (int e = 0; e < R; ++e) s += k.a[c][e] * k.a[c][e];
    part[c] = s;
  }
  T nrm = T(0);
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) {
    const T sc = wsum(part[c]);
    nrm = lane == c ? sc : nrm;
  }

  T off = T(0);
  if (intraI || intraJ) {
#pragma unroll 1
    for (int r = 0; r < B - 1; ++r) {
      jround<T, B, R, RV, 1>(k, nrm, intraI, intraJ, tol, off);
      shift_intra(k, nrm);
    }
  }
  if (cJ >= 0) {
#pragma unroll 1
    for (int r = 0; r < B; ++r) {
      jround<T, B, R, RV, 0>(k, nrm, true, true, tol, off);
      shift_cross(k, nrm);
    }
  }
  // recompute the store addresses here instead of keeping the load
  // addresses live through the rounds
  asm volatile("" : "+v"(cI), "+v"(cJ), "+v"(mt), "+v"(nt));
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) {
    const int col = c < B ? cI + c : cJ + (c - B);
    const bool ok = (c < B || cJ >= 0) && col < nt;
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const int r = lane + 64 * e;
      bw.st(ok && r < mt ? r + col * mt : -1, k.a[c][e]);
    }
#pragma unroll
    for (int e = 0; e < RV; ++e) {
      const int r = lane + 64 * e;
      bv.st(ok && r < nt ? r + col * nt : -1, k.v[c][e]);
    }
  }
  return wmax(off);
}

// Pair i of round r of the round-robin schedule of P players (circle method):
// every pair of 0..P-1 (P even) appears once in rounds 0..P-2.
__host__ __device__ constexpr int rr_p(int P, int r, int i) {
  return i == 0 ? 0 : 1 + (i - 1 + r) % (P - 1);
}
__host__ __device__ constexpr int rr_q(int P, int r, int i) { return 1 + (P - 2 - i + r) % (P - 1); }

// A team: the nw waves of a workgroup that factor one matrix together
// (nw = 1: a wave alone).  w is the wave's index in the team.
struct Team {
  int w, nw;
  __device__ __forceinline__ void sync() const {
    if (nw > 1) {
      __syncthreads();
    } else {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
};

template <typename T>
__device__ __forceinline__ T team_max(T v, Team t, T* red) {
  v = wmax(v);
  if (t.nw == 1) return v;
  if ((threadIdx.x & 63) == 0) red[t.w] = v;
  __syncthreads();
  T r = red[0];
  for (int i = 1; i < t.nw; ++i) r = fmax(r, red[i]);
  __syncthreads();
  return r;
}

// Sweeps of one matrix; returns the sweeps used or -1.  Each sweep is nbe-1
// rounds of the round-robin ordering of the nbe (nb rounded up to even)
// column blocks: the nbe/2 block pairs of a round are disjoint, so the
// team's waves take them in turn.  Round 0 also runs every block's internal
// pairs; a block paired with the dummy block nb runs alone.
template <typename T, int B, int R, int RV>
__device__ __forceinline__ int bj_sweeps(T* W, T* WV, int mt, int nt, Team t, T* red) {
  const T tol = Num<T>::eps * sqrt(T(mt));
  const int nb = (nt + B - 1) / B;
  const int nbe = nb + (nb & 1);
  for (int sw = 0; sw < BJ_MAXSWEEP; ++sw) {
    T off = T(0);
    for (int r = 0; r < nbe - 1; ++r) {
      for (int i = t.w; i < nbe / 2; i += t.nw) {
        int I = rr_p(nbe, r, i), J = rr_q(nbe, r, i);
        if (I > J) {
          const int x = I;
          I = J;
          J = x;
        }
        const bool alone = J >= nb;
        if (alone && r != 0) continue;
        off = fmax(off, block_step<T, B, R, RV>(W, WV, mt, nt, I * B, alone ? -1 : J * B, r == 0,
                                                r == 0 && !alone, tol));
      }
      t.sync();
    }
    off = team_max(off, t, red);
    if (!(off > tol)) return sw + 1;
  }
  return -1;
}

// Copies A (transposed when wide) into W scaled by a power of two so that
// max |a| is in [0.5, 1), and sets WV = I.  Returns the scale.
template <typename T>
__device__ T bj_load(const BatchItem& it, const T* A, T* W, T* WV, int mt, int nt, Team t,
                     T* red) {
  const int lane = threadIdx.x & 63;
  const bool wide = it.m < it.n;
  const T* src = A + it.a;
  T amax = T(0);
  for (int j = t.w; j < nt; j += t.nw)
    for (int r = lane; r < mt; r += 64) {
      const T x = wide ? src[j + (int64_t)r * it.m] : src[r + (int64_t)j * it.m];
      amax = fmax(amax, fabs(x));
    }
  amax = team_max(amax, t, red);
  T scale = T(1);
  if (amax > T(0) && isfinite(amax)) {
    int ex;
    (void)frexp(amax, &ex);
    scale = ldexp(T(1), -ex);
  }
  for (int j = t.w; j < nt; j += t.nw) {
    for (int r = lane; r < mt; r += 64) {
      const T x = wide ? src[j + (int64_t)r * it.m] : src[r + (int64_t)j * it.m];
      W[r + (int64_t)j * mt] = x * scale;
    }
    for (int r = lane; r < nt; r += 64) WV[r + (int64_t)j * nt] = r == j ? T(1) : T(0);
  }
  t.sync();
  return scale;
}

// sigma, ranks (ascending), outputs in order, completion of zero columns.
template <typename T>
__device__ void bj_finish(const BatchItem& it, const BjArgs<T>& g, T* W, T* WV, int mt, int nt,
                          T scale, T* sig, int* rk, Team t) {
  const int lane = threadIdx.x & 63;
  const int tt = t.w * 64 + lane;
  const bool wide = it.m < it.n;
  for (int j = t.w; j < nt; j += t.nw) {
    T s = T(0);
    for (int r = lane; r < mt; r += 64) {
      const T x = W[r + (int64_t)j * mt];
      s += x * x;
    }
    s = wsum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  t.sync();
  for (int j = tt; j < nt; j += 64 * t.nw) {
    const T sj = sig[j];
    int c = 0;
    for (int i = 0; i < nt; ++i) {
      const T si = sig[i];
      c += (si < sj || (si == sj && i < j)) ? 1 : 0;
    }
    rk[j] = c;
  }
  int nzero = 0;
  for (int i = 0; i < nt; ++i) nzero += sig[i] == T(0) ? 1 : 0;
  t.sync();
  // Y receives the normalised working columns (mt rows), Z the columns of V (nt rows)
  T* Y = wide ? g.V + it.v : g.U + it.u;
  T* Z = wide ? g.U + it.u : g.V + it.v;
  for (int j = t.w; j < nt; j += t.nw) {
    const int rj = rk[j];
    const T sj = sig[j];
    const T inv = sj > T(0) ? T(1) / sj : T(0);
    for (int r = lane; r < mt; r += 64) Y[r + (int64_t)rj * mt] = W[r + (int64_t)j * mt] * inv;
    for (int r = lane; r < nt; r += 64) Z[r + (int64_t)rj * nt] = WV[r + (int64_t)j * nt];
    if (lane == 0) g.S[it.s + rj] = sj / scale;
  }
  t.sync();
  // zero singular values sort first: columns 0..nzero-1 of Y are zero
  if (t.w == 0) {
    for (int z = 0; z < nzero; ++z) {
      for (int cand = 0; cand < mt; ++cand) {
        T x[BJ_MAXDIM / 64];
#pragma unroll
        for (int e = 0; e < BJ_MAXDIM / 64; ++e) x[e] = (lane + 64 * e == cand) ? T(1) : T(0);
        for (int pass = 0; pass < 2; ++pass) {
          for (int c = 0; c < nt; ++c) {
            if (c >= z && c < nzero) continue;
            T d = T(0);
#pragma unroll
            for (int e = 0; e < BJ_MAXDIM / 64; ++e) {
              const int r = lane + 64 * e;
              if (r < mt) d += Y[r + (int64_t)c * mt] * x[e];
            }
            d = wsum(d);
#pragma unroll
            for (int e = 0; e < BJ_MAXDIM / 64; ++e) {
              const int r = lane + 64 * e;
              if (r < mt) x[e] -= d * Y[r + (int64_t)c * mt];
            }
          }
        }
        T n2 = T(0);
#pragma unroll
        for (int e = 0; e < BJ_MAXDIM / 64; ++e) n2 += x[e] * x[e];
        n2 = wsum(n2);
        if (n2 > T(0.25)) {
          const T inv = T(1) / sqrt(n2);
#pragma unroll
          for (int e = 0; e < BJ_MAXDIM / 64; ++e) {
            const int r = lane + 64 * e;
            if (r < mt) Y[r + (int64_t)z * mt] = x[e] * inv;
          }
          break;
        }
      }
    }
  }
  t.sync();
}

template <typename T, int B, int R, int RV>
__device__ __forceinline__ void bj_matrix(const BatchItem& it, const BjArgs<T>& g, int k,
                                          Team t, T* sig, int* rk, T* red) {
  const int mt = it.m > it.n ? it.m : it.n, nt = it.m > it.n ? it.n : it.m;
  T* W = g.W + it.wa;
  T* WV = g.WV + it.wv;
  const T scale = bj_load(it, g.A, W, WV, mt, nt, t, red);
  const int sweeps = bj_sweeps<T, B, R, RV>(W, WV, mt, nt, t, red);
  bj_finish(it, g, W, WV, mt, nt, scale, sig, rk, t);
  if (g.info && threadIdx.x % (64 * t.nw) == 0) g.info[k] = sweeps;
}

__device__ __forceinline__ int64_t uni64(int64_t x) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(x & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
  return ((int64_t)hi << 32) | (int64_t)(unsigned)lo;
}

// The descriptor as wave-uniform (scalar) values.
__device__ __forceinline__ BatchItem uniform_item(const BatchItem& x) {
  BatchItem y;
  y.m = __builtin_amdgcn_readfirstlane(x.m);
  y.n = __builtin_amdgcn_readfirstlane(x.n);
  y.a = uni64(x.a);
  y.s = uni64(x.s);
  y.u = uni64(x.u);
  y.v = uni64(x.v);
  y.wa = uni64(x.wa);
  y.wv = uni64(x.wv);
  return y;
}

__device__ __forceinline__ int pop_wave(int* counter) {
  int idx = 0;
  if ((threadIdx.x & 63) == 0) idx = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(idx);
}

// Persistent kernel, BJ_WGW waves per workgroup.  While the queue head holds
// matrices with min(m, n) >= coop_nt, the workgroup's waves factor each one
// together; from the first smaller matrix on (the queue is sorted by work)
// every wave takes matrices on its own.
// B12 / B34: block width for matrices with up to 128 / up to 256 rows.
// MINW: resident waves per SIMD the register allocation must allow.
template <typename T, int B12, int B34, int MINW>
__global__ __launch_bounds__(64 * BJ_WGW, MINW) void bj_kernel(BjArgs<T> g) {
  __shared__ T sig[BJ_WGW][BJ_MAXDIM];
  __shared__ int rk[BJ_WGW][BJ_MAXDIM];
  __shared__ T red[BJ_WGW];
  __shared__ int s_idx;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  Team t{w, BJ_WGW};
  bool coop = true;
  for (;;) {
    int idx;
    if (coop) {
      if (threadIdx.x == 0) s_idx = atomicAdd(g.counter, 1);
      __syncthreads();
      idx = __builtin_amdgcn_readfirstlane(s_idx);
      __syncthreads();
    } else {
      idx = pop_wave(g.counter);
    }
    if (idx >= g.count) return;
    int k = __builtin_amdgcn_readfirstlane(g.order[idx]);
    BatchItem it = uniform_item(g.items[k]);
    if (coop && (it.m < it.n ? it.m : it.n) < g.coop_nt) {
      coop = false;
      t = Team{0, 1};
      if (w != 0) {  // wave 0 keeps this matrix, the others draw their own
        idx = pop_wave(g.counter);
        if (idx >= g.count) return;
        k = __builtin_amdgcn_readfirstlane(g.order[idx]);
        it = uniform_item(g.items[k]);
      }
    }
    T* sg = sig[coop ? 0 : w];
    int* rkk = rk[coop ? 0 : w];
    T* rd = coop ? red : red + w;
    const int mt = it.m > it.n ? it.m : it.n, nt = it.m > it.n ? it.n : it.m;
    const int R = (mt + 63) / 64, RV = (nt + 63) / 64;
#define BJ_CASE(R_, RV_)                                                  \
  if (R == R_ && RV == RV_) {                                             \
    bj_matrix<T, (R_ <= 2 ? B12 : B34), R_, RV_>(it, g, k, t, sg, rkk, rd); \
    continue;                                                             \
  }
    BJ_CASE(1, 1)
    BJ_CASE(2, 1)
    BJ_CASE(2, 2)
    BJ_CASE(3, 1)
    BJ_CASE(3, 2)
    BJ_CASE(3, 3)
    BJ_CASE(4, 1)
    BJ_CASE(4, 2)
    BJ_CASE(4, 3)
    BJ_CASE(4, 4)
#undef BJ_CASE
  }
}

__global__ void k_fill_i32(int* p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct BjVariant {
  int b12, b34;
  const void* f64;
  const void* f32;
};

#define BJ_VARIANT(B12, B34, W64, W32) \
  {B12, B34, (const void*)&bj_kernel<double, B12, B34, W64>, \
   (const void*)&bj_kernel<float, B12, B34, W32>}
static const BjVariant kVariants[] = {
    BJ_VARIANT(8, 4, 2, 3),
    BJ_VARIANT(4, 4, 2, 3),
    BJ_VARIANT(4, 2, 3, 4),
    BJ_VARIANT(2, 2, 3, 4),
};
#undef BJ_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
static const int kWgsPerCu[] = {2, 3};
// first entry: the threshold used while the variants are compared
static const int kCoopNt[] = {32, 1, 8, 16, 64, BJ_MAXDIM + 1};

// A launch configuration: kernel variant, resident workgroups per CU, and
// the smallest min(m, n) a whole workgroup factors together.
struct BjConfig {
  int variant = -1, wgs = 0, coop_nt = 0;
};

template <typename T>
static int bj_launch(const BjConfig& c, int cus, BjArgs<T> a, hipStream_t st) {
  const BjVariant& v = kVariants[c.variant];
  const void* f = sizeof(T) == 8 ? v.f64 : v.f32;
  a.coop_nt = c.coop_nt;
  void* params[] = {&a};
  DDC_HIP(hipLaunchKernel(f, dim3(cus * c.wgs), dim3(64 * BJ_WGW), params, 0, st));
  return DDC_OK;
}

// tuned configurations by (dtype, log2 count, log2 work)
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, BjConfig> g_tune_cache;

}  // namespace ddc

struct svd_gpu_batch {
  int count = 0, dtype = 0, cus = 0, dev = 0;
  std::vector<int> m, n;
  std::vector<long long> offA, offS, offU, offV;
  std::vector<int> order;
  double work = 0.0;
  ddc::BatchItem* d_items = nullptr;
  int* d_order = nullptr;   // processing order
  int* d_counter = nullptr;  // [0] queue head
  void* d_work = nullptr;
  size_t work_elems_a = 0, work_elems_v = 0;
  ddc::BjConfig cfg;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* gA = nullptr;
  void *gS = nullptr, *gU = nullptr, *gV = nullptr;
  int* ginfo = nullptr;
  std::mutex mu;
};

namespace ddc {

template <typename T>
static BjArgs<T> bj_args(svd_gpu_batch* p, const void* dA, void* dS, void* dU, void* dV,
                         int* info) {
  BjArgs<T> a;
  a.count = p->count;
  a.items = p->d_items;
  a.order = p->d_order;
  a.counter = p->d_counter;
  a.A = (const T*)dA;
  a.S = (T*)dS;
  a.U = (T*)dU;
  a.V = (T*)dV;
  a.W = (T*)p->d_work;
  a.WV = (T*)p->d_work + p->work_elems_a;
  a.info = info;
  return a;
}

static int bj_enqueue(svd_gpu_batch* p, const BjConfig& c, const void* dA, void* dS, void* dU,
                      void* dV, int* info, hipStream_t st) {
  int* ctr = p->d_counter;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, ctr, 0, 1);
  DDC_LAUNCHED();
  if (p->dtype == SVD_GPU_F64)
    return bj_launch<double>(c, p->cus, bj_args<double>(p, dA, dS, dU, dV, info), st);
  return bj_launch<float>(c, p->cus, bj_args<float>(p, dA, dS, dU, dV, info), st);
}

static std::tuple<int, int, int> tune_key(const svd_gpu_batch* p) {
  return {p->dtype, (int)std::lround(std::log2((double)p->count)),
          (int)std::lround(std::log2(p->work + 1.0))};
}

// Milliseconds of one launch of configuration c on the whole batch (best of 2).
static int bj_time(svd_gpu_batch* p, const BjConfig& c, const void* dA, void* dS, void* dU,
                   void* dV, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float& ms) {
  ms = 1e30f;
  for (int rep = 0; rep < 2; ++rep) {
    DDC_HIP(hipEventRecord(e0, st));
    DDC_TRY(bj_enqueue(p, c, dA, dS, dU, dV, nullptr, st));
    DDC_HIP(hipEventRecord(e1, st));
    DDC_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    DDC_HIP(hipEventElapsedTime(&t, e0, e1));
    ms = std::min(ms, t);
  }
  return DDC_OK;
}

// The autotuner: times (variant, workgroups per CU) with a middle
// cooperation threshold, then the thresholds for the best pair, each on the
// whole batch.  (A half-size sample was tried: it keeps the largest
// matrices, whose latency floors the makespan, but only half of the bulk,
// and ranked configurations 25% apart in the wrong order.)
static int bj_tune(svd_gpu_batch* p, const void* dA, void* dS, void* dU, void* dV,
                   hipStream_t st) {
  hipEvent_t e0, e1;
  DDC_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return DDC_E_HIP;
  }
  float best = 1e30f;
  BjConfig bc;
  int rc = DDC_OK;
  for (int v = 0; v < kNumVariants && !rc; ++v)
    for (int wg : kWgsPerCu) {
      BjConfig c;
      c.variant = v;
      c.wgs = wg;
      c.coop_nt = kCoopNt[0];
      float ms;
      if ((rc = bj_time(p, c, dA, dS, dU, dV, st, e0, e1, ms))) break;
      if (ms < best) {
        best = ms;
        bc = c;
      }
    }
  for (int ct : kCoopNt) {
    if (rc || ct == kCoopNt[0]) continue;
    BjConfig c = bc;
    c.coop_nt = ct;
    float ms;
    if ((rc = bj_time(p, c, dA, dS, dU, dV, st, e0, e1, ms))) break;
    if (ms < best) {
      best = ms;
      bc = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  p->cfg = bc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[tune_key(p)] = bc;
  return DDC_OK;
}

static void bj_free(svd_gpu_batch* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->d_items) (void)hipFree(p->d_items);
  if (p->d_order) (void)hipFree(p->d_order);
  if (p->d_counter) (void)hipFree(p->d_counter);
  if (p->d_work) (void)hipFree(p->d_work);
  p->exec = nullptr;
  p->graph = nullptr;
  p->cap = nullptr;
  p->d_items = nullptr;
  p->d_order = nullptr;
  p->d_counter = nullptr;
  p->d_work = nullptr;
}

// Items (element offsets into the packed buffers and into the work area)
// and costs of a batch, in batch order; off* (count + 1 entries) may be
// null.  wa / wv: elements of the working copies of A and V.
static void bj_layout(int count, const int* m, const int* n, BatchItem* items, double* cost,
                      long long* offA, long long* offS, long long* offU, long long* offV,
                      int64_t& wa, int64_t& wv) {
  long long oa = 0, os = 0, ou = 0, ov = 0;
  wa = wv = 0;
  for (int k = 0; k < count; ++k) {
    const int mk = m[k], nk = n[k], mn = std::min(mk, nk), mt = std::max(mk, nk);
    if (offA) {
      offA[k] = oa;
      offS[k] = os;
      offU[k] = ou;
      offV[k] = ov;
    }
    BatchItem& it = items[k];
    it.m = mk;
    it.n = nk;
    it.a = oa;
    it.s = os;
    it.u = ou;
    it.v = ov;
    it.wa = wa;
    it.wv = wv;
    oa += (long long)mk * nk;
    os += mn;
    ou += (long long)mk * mn;
    ov += (long long)nk * mn;
    wa += (int64_t)mt * mn;
    wv += (int64_t)mn * mn;
    cost[k] = (double)mn * mn * mt;
  }
  if (offA) {
    offA[count] = oa;
    offS[count] = os;
    offU[count] = ou;
    offV[count] = ov;
  }
}

static int bj_plan(svd_gpu_batch* p, int count, const int* m, const int* n, int dtype) {
  const size_t es = dtype == SVD_GPU_F64 ? 8 : 4;
  p->count = count;
  p->dtype = dtype;
  p->m.assign(m, m + count);
  p->n.assign(n, n + count);
  p->offA.assign(count + 1, 0);
  p->offS.assign(count + 1, 0);
  p->offU.assign(count + 1, 0);
  p->offV.assign(count + 1, 0);
  std::vector<BatchItem> items(count);
  std::vector<double> cost(count);
  int64_t wa = 0, wv = 0;
  bj_layout(count, m, n, items.data(), cost.data(), p->offA.data(), p->offS.data(),
            p->offU.data(), p->offV.data(), wa, wv);
  for (int k = 0; k < count; ++k) p->work += cost[k];
  p->work_elems_a = (size_t)round_up(wa, 64);
  p->work_elems_v = (size_t)round_up(wv, 64);
  p->order.resize(count);
  std::iota(p->order.begin(), p->order.end(), 0);
  std::stable_sort(p->order.begin(), p->order.end(),
                   [&](int x, int y) { return cost[x] > cost[y]; });

  DDC_HIP(hipGetDevice(&p->dev));
  DDC_HIP(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->dev));
  if (p->cus < 1) p->cus = 1;
  const size_t order_len = p->order.size();
  if (hipMalloc(&p->d_items, sizeof(BatchItem) * count) != hipSuccess ||
      hipMalloc(&p->d_order, sizeof(int) * order_len) != hipSuccess ||
      hipMalloc(&p->d_counter, sizeof(int) * 64) != hipSuccess ||
      hipMalloc(&p->d_work, es * (p->work_elems_a + p->work_elems_v)) != hipSuccess) {
    (void)hipGetLastError();
    report_error("device allocation for the batch failed");
    return DDC_E_NOMEM;
  }
  std::vector<int> ord(p->order);
  DDC_HIP(hipMemcpy(p->d_items, items.data(), sizeof(BatchItem) * count, hipMemcpyHostToDevice));
  DDC_HIP(hipMemcpy(p->d_order, ord.data(), sizeof(int) * order_len, hipMemcpyHostToDevice));
  DDC_HIP(hipMemset(p->d_counter, 0, sizeof(int) * 64));
  DDC_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto f = g_tune_cache.find(tune_key(p));
    if (f != g_tune_cache.end()) p->cfg = f->second;
  }
  return DDC_OK;
}

static int bj_run(svd_gpu_batch* p, const void* dA, void* dS, void* dU, void* dV, int* info,
                  hipStream_t st) {
  if (p->cfg.variant < 0) DDC_TRY(bj_tune(p, dA, dS, dU, dV, st));
  const bool same = p->exec && p->gA == dA && p->gS == dS && p->gU == dU && p->gV == dV &&
                    p->ginfo == info;
  if (!same) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    DDC_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
    const int rc = bj_enqueue(p, p->cfg, dA, dS, dU, dV, info, p->cap);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    DDC_HIP(e);
    p->graph = gr;
    DDC_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    p->gA = dA;
    p->gS = dS;
    p->gU = dU;
    p->gV = dV;
    p->ginfo = info;
  }
  DDC_HIP(hipGraphLaunch(p->exec, st));
  return DDC_OK;
}

// ---------------------------------------------------------------------------
// The variant that allocates nothing: every buffer beyond the caller's inputs
// and outputs (matrix descriptors, processing order, queue head, working
// copies) lives in one file-scope device array, and the host-side
// descriptors in file-scope host arrays.  Each GPU holds its own copy of the
// device array (its address is looked up once per device); calls on one GPU
// share it, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t BJ_NA_BYTES = (size_t)32 << 20;  // device scratch per GPU
constexpr int BJ_NA_MAXCOUNT = 4096;              // matrices per call
constexpr int BJ_NA_MAXDEV = 64;
// fixed launch configuration: variant 1 (column blocks of 4), two resident
// workgroups per CU, whole-workgroup factorization from min(m, n) = 32 on
static const BjConfig kNoallocConfig = {1, 2, 32};

static __device__ __align__(256) char scratch[BJ_NA_BYTES];

namespace {
std::mutex g_na_mu;
char* g_na_ptr[BJ_NA_MAXDEV] = {};
int g_na_cus[BJ_NA_MAXDEV] = {};
BatchItem g_na_items[BJ_NA_MAXCOUNT];
double g_na_cost[BJ_NA_MAXCOUNT];
int g_na_order[BJ_NA_MAXCOUNT];
}  // namespace

static size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }

static int bj_noalloc(int count, const int* m, const int* n, int dtype, const void* dA, void* dS,
                      void* dU, void* dV, int* info, hipStream_t st) {
  int dev = 0;
  DDC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= BJ_NA_MAXDEV) {
    report_error("device index beyond the no-alloc table");
    return DDC_E_ARG;
  }
  if (!g_na_ptr[dev]) {
    // set-up, once per device with that device current: the device address
    // of this GPU's copy of the scratch array (the array's name in host
    // code is only a host-side handle)
    void* ptr = nullptr;
    DDC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    int cus = 0;
    DDC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    g_na_cus[dev] = cus > 0 ? cus : 1;
    g_na_ptr[dev] = (char*)ptr;
  }
  char* base = g_na_ptr[dev];
  const size_t es = dtype == SVD_GPU_F64 ? 8 : 4;
  int64_t wa = 0, wv = 0;
  bj_layout(count, m, n, g_na_items, g_na_cost, nullptr, nullptr, nullptr, nullptr, wa, wv);
  const size_t b_items = na_align(sizeof(BatchItem) * count);
  const size_t b_order = na_align(sizeof(int) * count);
  const size_t b_ctr = 256;
  const size_t ea = (size_t)round_up(wa, 64), ev = (size_t)round_up(wv, 64);
  const size_t need = b_items + b_order + b_ctr + na_align(es * (ea + ev));
  if (need > BJ_NA_BYTES) {
    char msg[160];
    snprintf(msg, sizeof(msg), "batch needs %zu bytes of scratch; the no-alloc limit is %zu",
             need, BJ_NA_BYTES);
    report_error(msg);
    return DDC_E_ARG;
  }
  for (int k = 0; k < count; ++k) g_na_order[k] = k;
  // largest first, ties by index (std::sort works in place)
  std::sort(g_na_order, g_na_order + count, [](int x, int y) {
    return g_na_cost[x] > g_na_cost[y] || (g_na_cost[x] == g_na_cost[y] && x < y);
  });
  BatchItem* d_items = (BatchItem*)base;
  int* d_order = (int*)(base + b_items);
  int* d_ctr = (int*)(base + b_items + b_order);
  char* d_work = base + b_items + b_order + b_ctr;
  DDC_HIP(hipMemcpyAsync(d_items, g_na_items, sizeof(BatchItem) * count, hipMemcpyHostToDevice,
                         st));
  DDC_HIP(hipMemcpyAsync(d_order, g_na_order, sizeof(int) * count, hipMemcpyHostToDevice, st));
  // the host arrays are reused by the next call
  DDC_HIP(hipStreamSynchronize(st));
  // the queue head starts at zero (a fill kernel; the array keeps the last
  // call's contents)
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d_ctr, 0, 1);
  DDC_LAUNCHED();
  if (dtype == SVD_GPU_F64) {
    BjArgs<double> a;
    a.count = count;
    a.items = d_items;
    a.order = d_order;
    a.counter = d_ctr;
    a.A = (const double*)dA;
    a.S = (double*)dS;
    a.U = (double*)dU;
    a.V = (double*)dV;
    a.W = (double*)d_work;
    a.WV = (double*)d_work + ea;
    a.info = info;
    return bj_launch<double>(kNoallocConfig, g_na_cus[dev], a, st);
  }
  BjArgs<float> a;
  a.count = count;
  a.items = d_items;
  a.order = d_order;
  a.counter = d_ctr;
  a.A = (const float*)dA;
  a.S = (float*)dS;
  a.U = (float*)dU;
  a.V = (float*)dV;
  a.W = (float*)d_work;
  a.WV = (float*)d_work + ea;
  a.info = info;
  return bj_launch<float>(kNoallocConfig, g_na_cus[dev], a, st);
}

}  // namespace ddc

using namespace ddc;

extern "C" {

int svd_gpu_batched_plan(svd_gpu_batch** plan, int count, const int* m, const int* n, int dtype) {
  clear_error();
  if (!plan) {
    report_error("null plan pointer");
    return DDC_E_ARG;
  }
  *plan = nullptr;
  if (count < 1 || !m || !n || (dtype != SVD_GPU_F64 && dtype != SVD_GPU_F32)) {
    report_error("invalid batch arguments");
    return DDC_E_ARG;
  }
  for (int k = 0; k < count; ++k)
    if (m[k] < 1 || n[k] < 1 || m[k] > BJ_MAXDIM || n[k] > BJ_MAXDIM) {
      report_error("batched matrices need 1 <= m, n <= 256");
      return DDC_E_ARG;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    report_error("no HIP device available");
    return DDC_E_NODEV;
  }
  svd_gpu_batch* p = new (std::nothrow) svd_gpu_batch;
  if (!p) return DDC_E_NOMEM;
  const int rc = bj_plan(p, count, m, n, dtype);
  if (rc) {
    bj_free(p);
    delete p;
    return rc;
  }
  *plan = p;
  return DDC_OK;
}

int svd_gpu_batched_offsets(const svd_gpu_batch* plan, long long* offA, long long* offS,
                            long long* offU, long long* offV) {
  if (!plan) return DDC_E_ARG;
  const size_t c = (size_t)plan->count + 1;
  if (offA) std::memcpy(offA, plan->offA.data(), c * sizeof(long long));
  if (offS) std::memcpy(offS, plan->offS.data(), c * sizeof(long long));
  if (offU) std::memcpy(offU, plan->offU.data(), c * sizeof(long long));
  if (offV) std::memcpy(offV, plan->offV.data(), c * sizeof(long long));
  return DDC_OK;
}

int svd_gpu_batched_tune(svd_gpu_batch* plan, const void* dA, void* dsigma, void* dU, void* dV,
                         void* stream) {
  clear_error();
  if (!plan || !dA || !dsigma || !dU || !dV) {
    report_error("invalid batch arguments");
    return DDC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(plan->mu);
  return bj_tune(plan, dA, dsigma, dU, dV, (hipStream_t)stream);
}

int svd_gpu_batched_set_config(svd_gpu_batch* plan, int variant, int wgs_per_cu,
                               int coop_min_n) {
  clear_error();
  if (!plan || variant < 0 || variant >= kNumVariants || wgs_per_cu < 1 || wgs_per_cu > 8 ||
      coop_min_n < 1) {
    report_error("invalid batch configuration");
    return DDC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(plan->mu);
  plan->cfg.variant = variant;
  plan->cfg.wgs = wgs_per_cu;
  plan->cfg.coop_nt = coop_min_n;
  plan->gA = nullptr;  // forces a new graph
  plan->ginfo = nullptr;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  plan->exec = nullptr;
  return DDC_OK;
}

int svd_gpu_batched_config(const svd_gpu_batch* plan, int* variant, int* wgs_per_cu,
                           int* coop_min_n, int* block_small, int* block_large) {
  if (!plan) return DDC_E_ARG;
  const BjConfig& c = plan->cfg;
  if (variant) *variant = c.variant;
  if (wgs_per_cu) *wgs_per_cu = c.wgs;
  if (coop_min_n) *coop_min_n = c.coop_nt;
  if (block_small) *block_small = c.variant >= 0 ? kVariants[c.variant].b12 : 0;
  if (block_large) *block_large = c.variant >= 0 ? kVariants[c.variant].b34 : 0;
  return kNumVariants;
}

int svd_gpu_batched_run(svd_gpu_batch* plan, const void* dA, void* dsigma, void* dU, void* dV,
                        int* dinfo, void* stream) {
  clear_error();
  if (!plan || !dA || !dsigma || !dU || !dV) {
    report_error("invalid batch arguments");
    return DDC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(plan->mu);
  return bj_run(plan, dA, dsigma, dU, dV, dinfo, (hipStream_t)stream);
}

int svd_gpu_batched_noalloc(int count, const int* m, const int* n, int dtype, const void* dA,
                            void* dsigma, void* dU, void* dV, int* dinfo, void* stream) {
  clear_error();
  if (count < 1 || count > BJ_NA_MAXCOUNT || !m || !n || !dA || !dsigma || !dU || !dV ||
      (dtype != SVD_GPU_F64 && dtype != SVD_GPU_F32)) {
    report_error("invalid batch arguments (1 <= count <= 4096 for the no-alloc entry point)");
    return DDC_E_ARG;
  }
  for (int k = 0; k < count; ++k)
    if (m[k] < 1 || n[k] < 1 || m[k] > BJ_MAXDIM || n[k] > BJ_MAXDIM) {
      report_error("batched matrices need 1 <= m, n <= 256");
      return DDC_E_ARG;
    }
  std::lock_guard<std::mutex> lk(g_na_mu);
  return bj_noalloc(count, m, n, dtype, dA, dsigma, dU, dV, dinfo, (hipStream_t)stream);
}

void svd_gpu_batched_destroy(svd_gpu_batch* plan) {
  if (!plan) return;
  {
    std::lock_guard<std::mutex> lk(plan->mu);
    bj_free(plan);
  }
  delete plan;
}

}  // extern "C"
