// twisted.hip -- per-singular-value vectors by twisted factorization, FP64.
//
// Replaces CalcRightSingularVectors and RighttoLeftSingularVectors
// (parallel-twisted.c:530-637).  The reference factors B^T B - sigma^2 I
// after forming B^T B (SquareB :290-316), which loses the small singular
// values, builds u = B v / sigma, which divides by zero for sigma = 0, and
// solves every singular value on its own, so repeated or clustered values
// get one and the same vector.
//
// Here every singular value gets one lane.  The matrix B^T B is kept in its
// factored form L D L^T with D = diag(a^2) and L = B^T diag(1/a), used only
// through d_i = a_i^2, (ld)_i = a_i b_i and (lld)_i = b_i^2 (so a zero
// diagonal entry needs no division).  The twisted factorization of
// L D L^T - lambda I comes from the stationary and progressive differential
// qd transforms; its smallest twist gamma_r gives the vector (N_r^T z = e_r)
// and the Rayleigh-quotient correction lambda += gamma_r / |z|^2.  Each
// iteration is kept inside a bracket certified by Sturm counts (negative
// pivots of the stationary transform), so the shift cannot drift to a
// neighbouring singular value.  On exit sigma holds the refined value.
//
// Left vectors are computed the same way from B B^T, which is the Gram
// matrix of the reversed bidiagonal B' = J B^T J (diagonal a reversed,
// super-diagonal b reversed); u = J z'.  The sign of u is then fixed so that
// u^T B v >= 0.
//
// Repeated and clustered singular values:
//  * split: a super-diagonal entry with |b_i| <= 16 eps sigma_max is set to
//    zero, which moves no singular value by more than that; B becomes block
//    diagonal and every recurrence restarts at a block boundary;
//  * isolate: the divide-and-conquer values are grouped into clusters whose
//    neighbours lie within 2 dl (dl = 8 n eps sigma_max); all lanes of a
//    cluster start from the cluster's common bracket and bisect it with
//    Sturm counts until their value is alone in it or the bracket is at the
//    rounding level.  Because the bisection is a function of the bracket and
//    the counts only, lanes whose values cannot be separated end with the
//    same bracket; they are then handed out to the blocks of B in block
//    order, using per-block counts, and each lane restricts its twist to its
//    block (so A = I, whose B = -I splits into n blocks, gets n different
//    unit vectors);
//  * Rayleigh-Ritz: lanes of one block whose refined values have a relative
//    gap in sigma^2 below TW_GTOL form a group.  For each group, the left and
//    right invariant subspaces are recomputed by three steps of inverse
//    iteration per member (pseudo-random start; own shift, or for a run of
//    members closer than TW_TIGHT eps one shift just below the run) and
//    orthonormalised
//    (Gram-Schmidt, twice); the k x k matrix M = U_c^T B V_c is diagonalised
//    by one-sided Jacobi and the bases are rotated by its singular vectors.
//    Windows hold at most 64 members; a wider group runs window by window
//    (one launch each), every window orthogonalised against the group's
//    earlier windows and, unless its values are at the rounding level of
//    zero, with its left basis formed as B V_c (independent left and right
//    iterates of an inseparable run wider than a window span different
//    subspaces of it).
#include <algorithm>
#include <vector>

#include "ddc_common.h"

namespace ddc {

constexpr double TW_EPS = 2.220446049250313e-16;
constexpr double TW_TINY = 2.2250738585072014e-308;
constexpr double TW_SPLIT = 16.0;  // split |b_i| <= TW_SPLIT eps sigma_max
constexpr double TW_GTOL = 1e-6;   // relative gap (in sigma^2) that forms a group
constexpr int TW_KMAX = 64;        // members per Rayleigh-Ritz window
constexpr double TW_TIGHT = 1024.0;  // members closer than this (eps, relative) share a shift
constexpr int TW_RRIT = 3;          // inverse-iteration steps per Rayleigh-Ritz member
constexpr int TW_CK = 64;          // checkpoint interval = LDS chunk
constexpr int TW_MAXBIS = 200;     // bisection steps per lane at most
constexpr double TW_CONV = 32.0;   // Rayleigh correction accepted below TW_CONV eps lam

enum { TW_UNRES = 1 };

__device__ __forceinline__ double guard(double x, double pivmin) {
  return (fabs(x) < pivmin) ? -pivmin : x;
}
// The quotient of a step is formed whether or not the step is a block
// boundary, and the boundary value selected afterwards: with no branch around
// the division, the operand loads of a run of steps can all be issued ahead
// of the dependent arithmetic (the empty asm, opaque to the optimiser but no
// scheduling barrier, keeps the compiler from moving the quotient back under
// a branch).  The selected values are those of the plain expressions.
__device__ __forceinline__ double keep(double x) {
  asm("" : "+v"(x));
  return x;
}
// stationary qd step: s_{i+1} from s_i (lld_i = 0 is a block boundary)
__device__ __forceinline__ double stat_next(double s, double dp, double lld, double lam) {
  const double q = keep(lld * s / dp);
  return lld == 0.0 ? -lam : q - lam;
}
// progressive qd step: p_i from p_{i+1}
__device__ __forceinline__ double prog_prev(double p, double d, double dm, double lld, double lam) {
  const double q = keep(d * p / dm);
  return lld == 0.0 ? d - lam : q - lam;
}

// qd data for B and its reversal, with negligible super-diagonal entries set
// to zero; split[i] = 1 where b_i was zeroed (i < n-1).
__global__ void tw_prep(int n, const double* a, const double* b, double tau, double* qd,
                        int* split) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* d0 = qd;
  double* ld0 = qd + n;
  double* lld0 = qd + 2 * (size_t)n;
  double* d1 = qd + 3 * (size_t)n;
  double* ld1 = qd + 4 * (size_t)n;
  double* lld1 = qd + 5 * (size_t)n;
  auto bb = [&](int k) { return fabs(b[k]) <= tau ? 0.0 : b[k]; };
  d0[i] = a[i] * a[i];
  const int r = n - 1 - i;
  d1[i] = a[r] * a[r];
  if (i < n - 1) {
    const double bi = bb(i), br = bb(r - 1);
    ld0[i] = a[i] * bi;
    lld0[i] = bi * bi;
    // reversed: a'_i = a[n-1-i], b'_i = b[n-2-i]
    ld1[i] = a[r] * br;
    lld1[i] = br * br;
    split[i] = bi == 0.0;
  } else {
    ld0[i] = lld0[i] = ld1[i] = lld1[i] = 0.0;
    split[i] = 1;
  }
}

// Inclusive scan of one value per thread over a 1024-thread work-group
// (MAXOP: maximum, else minimum).
template <bool MAXOP>
__device__ int wg_scan(int v, int* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int w = t >= o ? sh[t - o] : v;
    __syncthreads();
    v = MAXOP ? max(v, w) : min(v, w);
    sh[t] = v;
    __syncthreads();
  }
  return v;
}

// One work-group of 1024 threads: block ranges of the split matrix
// (bs[i] = first index of i's block, be[i] = one past its last) and the
// clusters of the sorted values (cs[j], ce[j] = first and last index of j's
// cluster: neighbours closer than 2 dl); cnt[0] = number of splits.
__global__ __launch_bounds__(1024) void tw_scan(int n, const int* split, const double* sig,
                                                double dl, int* bs, int* be, int* cs, int* ce,
                                                int* cnt) {
  __shared__ int sh[1024];
  __shared__ int carry[2];
  const int t = threadIdx.x;
  int nsplit = 0;
  if (t == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + t;
    const bool in = i < n;
    const bool bstart = in && (i == 0 || split[i - 1]);
    const bool chead = in && (i == 0 || sig[i] - sig[i - 1] > 2.0 * dl);
    if (in && i < n - 1) nsplit += split[i];
    int v = wg_scan<true>(bstart ? i : 0, sh);
    v = max(v, carry[0]);
    int w = wg_scan<true>(chead ? i : 0, sh);
    w = max(w, carry[1]);
    if (in) {
      bs[i] = v;
      cs[i] = w;
    }
    __syncthreads();
    if (t == 1023) {
      carry[0] = v;
      carry[1] = w;
    }
    __syncthreads();
  }
  if (t == 0) carry[0] = carry[1] = n;
  __syncthreads();
  const int nchunk = (n + 1023) / 1024;
  for (int c = nchunk - 1; c >= 0; --c) {
    // reversed order inside the chunk so that an inclusive min-scan runs
    // from the end
    const int i = c * 1024 + (1023 - t);
    const bool in = i < n;
    const bool bend = in && (i == n - 1 || split[i]);
    const bool ctail = in && (i == n - 1 || sig[i + 1] - sig[i] > 2.0 * dl);
    int v = wg_scan<false>(bend ? i + 1 : n, sh);
    v = min(v, carry[0]);
    int w = wg_scan<false>(ctail ? i : n - 1, sh);
    w = min(w, carry[1]);
    if (in) {
      be[i] = v;
      ce[i] = w;
    }
    __syncthreads();
    if (t == 1023) {
      carry[0] = v;
      carry[1] = w;
    }
    __syncthreads();
  }
  // number of splits
  sh[t] = nsplit;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) cnt[0] = sh[0];
}

struct IsoArgs {
  int n;
  const double* qd;   // rep 0: d, ld, lld
  const double* sig;  // divide-and-conquer values (ascending)
  const int *bs, *be, *cs, *ce;
  const int* cnt;     // cnt[0] = number of splits
  double smax, dl;
  double* lam_lo;     // bracket of the lane's lambda = sigma^2
  double* lam_hi;
  int* r0;            // twist range [r0, r1) (rep 0 indices)
  int* r1;
  int* flags;
};

// Negative pivots of the stationary transform of L D L^T - x I at two shifts
// x0 <= x1 over [0, n) (x <= 0 counts nothing: B^T B is semi-definite).
// The loop index is uniform, so the qd loads are scalar.
__device__ __forceinline__ void sturm2(const double* d, const double* lld, int n, double x0,
                                       double x1, double pivmin, int& c0, int& c1) {
  double s0 = -x0, s1 = -x1;
  int k0 = 0, k1 = 0;
  for (int i = 0; i < n; ++i) {
    const double di = d[i], li = lld[i];
    const double p0 = guard(di + s0, pivmin), p1 = guard(di + s1, pivmin);
    k0 += p0 < 0.0;
    k1 += p1 < 0.0;
    s0 = stat_next(s0, p0, li, x0);
    s1 = stat_next(s1, p1, li, x1);
  }
  c0 = x0 > 0.0 ? k0 : 0;
  c1 = x1 > 0.0 ? k1 : 0;
}

// The same at one shift.
__device__ __forceinline__ int sturm1(const double* d, const double* lld, int n, double x,
                                      double pivmin) {
  double s = -x;
  int k = 0;
  for (int i = 0; i < n; ++i) {
    const double p = guard(d[i] + s, pivmin);
    k += p < 0.0;
    s = stat_next(s, p, lld[i], x);
  }
  return x > 0.0 ? k : 0;
}

// One lane per singular value (rep 0): the bracket that isolates the lane's
// eigenvalue of B^T B, or the rounding-level bracket of an inseparable
// cluster, and the block of B the lane's vector lives in.
__global__ __launch_bounds__(256) void tw_isolate(IsoArgs A) {
  const int n = A.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < n;
  const int jj = live ? j : n - 1;
  const double* d = A.qd;
  const double* lld = A.qd + 2 * (size_t)n;
  const double smax = A.smax, dl = A.dl;
  const double pivmin = TW_TINY * fmax(1.0, smax * smax);
  const double top = (smax + dl) * (smax + dl) * (1.0 + 1e-8);
  const int c_s = A.cs[jj], c_e = A.ce[jj];
  double lo = fmax(0.0, A.sig[c_s] - dl);
  lo *= lo;
  double hi = (A.sig[c_e] + dl) * (A.sig[c_e] + dl);
  int clo, chi;
  sturm2(d, lld, n, lo, hi, pivmin, clo, chi);
  // the cluster's common bracket must hold its values; widen it if the
  // values were off by more than dl
  if (clo > c_s) {
    lo = 0.0;
    clo = 0;
  }
  if (chi < c_e + 1) {
    hi = top;
    chi = n;
  }
  bool need = live && !(clo == jj && chi == jj + 1);
  bool unres = false;
  int it = 0;
  while (__any(need)) {
    double mid = 0.0;
    if (need) {
      if (hi - lo <= 4.0 * TW_EPS * hi || hi <= 4.0 * pivmin || it >= TW_MAXBIS) {
        unres = true;
        need = false;
      } else {
        mid = (lo > 0.0 && hi > 2.0 * lo) ? sqrt(lo) * sqrt(hi) : 0.5 * (lo + hi);
        ++it;
      }
    }
    if (!__any(need)) break;
    const int cm = sturm1(d, lld, n, need ? mid : 0.0, pivmin);
    if (need) {
      if (cm <= jj) {
        lo = mid;
        clo = cm;
      } else {
        hi = mid;
        chi = cm;
      }
      if (clo == jj && chi == jj + 1) need = false;
    }
  }
  int b0 = 0, b1 = n;
  // a lane that had to bisect is given a block: the t-th value of the
  // bracket, counting block by block
  const bool walk = A.cnt[0] > 0 && it > 0 && live;
  if (__any(walk)) {
    const int t = jj - clo;
    double s0 = -lo, s1 = -hi;
    int k0 = 0, k1 = 0, cum = 0;
    bool found = !walk;
    for (int i = 0; i < n; ++i) {
      if (A.bs[i] == i) {
        s0 = -lo;
        s1 = -hi;
        k0 = k1 = 0;
      }
      const double di = d[i], li = lld[i];
      const double p0 = guard(di + s0, pivmin), p1 = guard(di + s1, pivmin);
      k0 += p0 < 0.0;
      k1 += p1 < 0.0;
      s0 = stat_next(s0, p0, li, lo);
      s1 = stat_next(s1, p1, li, hi);
      const int e = A.be[i];
      if (e == i + 1) {
        const int mb = (hi > 0.0 ? k1 : 0) - (lo > 0.0 ? k0 : 0);
        if (!found && cum + mb > t) {
          found = true;
          b0 = A.bs[i];
          b1 = e;
        }
        cum += mb;
      }
    }
  }
  if (!live) return;
  A.lam_lo[j] = lo;
  A.lam_hi[j] = hi;
  A.r0[j] = b0;
  A.r1[j] = b1;
  A.flags[j] = unres ? TW_UNRES : 0;
}

struct TwArgs {
  int n;
  const double* qd;      // [2][3][n]: rep 0 = B, rep 1 = reversed; d, ld, lld
  const double* sig_in;  // D&C values
  double* sig_out;       // refined values
  double* Vt;            // v_i[j] at j*ldv + i
  double* Ut;            // u_i[j] at j*ldv + i
  int ldv;
  double* Sck;           // checkpoints of s every TW_CK steps [nck][T]
  double* Pck;           // checkpoints of p every TW_CK steps [nck][T]
  int T;                 // checkpoint stride (2n)
  double smax;
  const double* lam_lo;  // per-lane bracket and twist range (tw_isolate)
  const double* lam_hi;
  const int* r0;
  const int* r1;
  const int* flags;
  const int* bs;         // block start of each index
  int* blk;              // out: block start of the lane's vector (rep 0)
};

// Two waves per group of 64 singular values, all lanes on the same
// representation; lane = one singular value, and lane l of both waves works
// on the same value (their per-value state is kept identical).  The lanes
// walk the index range in lockstep, chunk by chunk; the chunk's qd entries
// are staged in LDS, and the recurrences that a pass needs in the opposite
// direction are recomputed for the chunk from checkpoints (every TW_CK
// steps) into a per-wave LDS buffer.  Nothing of size n per lane is stored
// except the vector itself.  The two waves run the independent recurrences
// of each pass side by side, on different SIMDs:
//   pass 1   wave 0: stationary transform, ascending (checkpoints of s,
//            Sturm count at lam); wave 1: progressive transform,
//            descending (checkpoints of p)
//   pass 2   per chunk, wave 0 recomputes s and wave 1 recomputes p; gamma
//            and its argmin over the chunk's two halves, among the indices
//            of the lane's twist range
//   pass 3   wave 0: z downward from the twist (recomputed p); wave 1: z
//            upward (recomputed s)
//
// Recurrences (L D L^T - lam I, d = a^2, ld = a b, lld = b^2; lld_i = 0
// restarts them, B being block diagonal there):
//   stationary  s_0 = -lam, D+_i = d_i + s_i, s_{i+1} = lld_i s_i / D+_i - lam
//   progressive p_{n-1} = d_{n-1} - lam, D-_{i+1} = lld_i + p_{i+1},
//               p_i = d_i p_{i+1} / D-_{i+1} - lam
//   twist       gamma_i = s_i + p_i + lam, r = argmin |gamma_i| (the largest
//               such i on ties)
//   vector      z_r = 1, z_{i+1} = -(ld_i / D-_{i+1}) z_i  (i >= r),
//               z_i = -(ld_i / D+_i) z_{i+1}  (i < r)
//   correction  lam += gamma_r / |z|^2   (Rayleigh quotient)
__global__ __launch_bounds__(128) void tw_solve(TwArgs a) {
  // staged qd of the chunk each wave works on: [wave][d, ld, lld][e]
  __shared__ double qs[2][3][TW_CK];
  __shared__ double buf[2][TW_CK][64];  // per-wave recomputed s or p, [e][lane]
  // exchange between the waves, [wave][lane]
  __shared__ double xg[2][64], xn[2][64];
  __shared__ int xr[2][64], xc[64];
  const int n = a.n;
  const int nbr = (n + 63) / 64;
  const int rep = ((int)blockIdx.x >= nbr) ? 1 : 0;
  // wave is uniform over the wave: chunk indices and lengths stay scalar,
  // so the full-chunk loops below are straight-line code
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = ((int)blockIdx.x - rep * nbr) * 64 + lane;
  const bool live = idx < n;
  const int li = live ? idx : 0;
  const double* d = a.qd + (size_t)rep * 3 * n;
  const int64_t T = a.T;
  const int col = rep * n + li;
  double* Sck = a.Sck + col;
  double* Pck = a.Pck + col;
  double* out = (rep ? a.Ut : a.Vt) + li;
  const int64_t ldv = a.ldv;
  auto oidx = [&](int j) -> int64_t { return (int64_t)(rep ? (n - 1 - j) : j) * ldv; };
  const int nck = (n + TW_CK - 1) / TW_CK;
  // both waves stage one chunk each (cw: this wave's chunk)
  auto stage = [&](int cw) {
    __syncthreads();
    const int c0 = cw * TW_CK, len = min(TW_CK, n - c0);
    if (lane < len) {
      qs[wave][0][lane] = d[c0 + lane];
      qs[wave][1][lane] = d[n + c0 + lane];
      qs[wave][2][lane] = d[2 * n + c0 + lane];
    }
    __syncthreads();
  };
  const double* qd_d = qs[wave][0];
  const double* qd_ld = qs[wave][1];
  const double* qd_lld = qs[wave][2];
  double (*bw)[64] = buf[wave];

  const double smax = a.smax;
  const double pivmin = TW_TINY * fmax(1.0, smax * smax);
  const double sig = live ? a.sig_in[idx] : 0.0;
  double lo = a.lam_lo[li], hi = a.lam_hi[li];
  const bool unres = (a.flags[li] & TW_UNRES) != 0;
  // twist range in this representation's indices
  const int g0 = rep ? n - a.r1[li] : a.r0[li];
  const int g1 = rep ? n - a.r0[li] : a.r1[li];
  const double dn1 = d[n - 1];

  double lam = fmin(fmax(sig * sig, lo), hi);
  // A value whose square is within the pivot guard (0 in particular) is
  // solved at a shift just below zero (eps^2 smax^2, at least 4 pivmin so
  // that it is not itself guarded), where the shifted Gram
  // matrix is positive definite: at a shift of exactly 0 a zero diagonal
  // entry makes the transforms' quotients 0/0 (taken as 0 by the guard
  // instead of their limit), every twist index ties and the vector is
  // wrong.  Its first vector is kept.  (Larger tiny values keep their own
  // shift: in a graded block they are resolved to high relative accuracy.)
  if (lam <= pivmin) lam = -fmax(TW_EPS * TW_EPS * smax * smax, 4.0 * pivmin);
  double nz2 = 1.0, corr = 0.0, prev = INFINITY;
  int rfin = g1 - 1;
  bool active = live;
  for (int it = 0; it < 10; ++it) {
    if (__ballot(active) == 0) break;  // the same in both waves
    // pass 1: stationary (wave 0, ascending) and progressive (wave 1,
    // descending) transforms with checkpoints
    int cnt = 0;
    {
      double s = -lam;           // wave 0
      double p = dn1 - lam;      // wave 1
      for (int c = 0; c < nck; ++c) {
        const int cw = wave == 0 ? c : nck - 1 - c;
        stage(cw);
        const int c0 = cw * TW_CK, len = min(TW_CK, n - c0);
        if (wave == 0) {
          if (active) Sck[(int64_t)cw * T] = s;
          auto step = [&](int e) {
            const double dp = guard(qd_d[e] + s, pivmin);
            cnt += (dp < 0.0);
            s = stat_next(s, dp, qd_lld[e], lam);  // past the last index: unused
          };
          if (len == TW_CK) {
#pragma unroll 8
            for (int e = 0; e < TW_CK; ++e) step(e);
          } else {
            for (int e = 0; e < len; ++e) step(e);
          }
        } else {
          auto step = [&](int e) {
            const double dm = guard(qd_lld[e] + p, pivmin);
            p = prog_prev(p, qd_d[e], dm, qd_lld[e], lam);
          };
          // the chunk's steps e < len with c0 + e < n - 1, descending
          const int etop = min(len, n - 1 - c0) - 1;
          if (etop == TW_CK - 1) {
#pragma unroll 8
            for (int e = TW_CK - 1; e >= 0; --e) step(e);
          } else {
            for (int e = etop; e >= 0; --e) step(e);
          }
          if (active) Pck[(int64_t)cw * T] = p;  // p at the chunk's first index
        }
      }
    }
    // pass 2: gamma_i = s_i + p_i + lam over each chunk, both recomputed
    int r = -1;  // no finite candidate yet
    double gbest = INFINITY;
    for (int c = 0; c < nck; ++c) {
      stage(c);
      const int c0 = c * TW_CK, len = min(TW_CK, n - c0);
      if (wave == 0) {
        double ss = Sck[(int64_t)c * T];
        auto step = [&](int e) {
          bw[e][lane] = ss;
          const double dp = guard(qd_d[e] + ss, pivmin);
          ss = stat_next(ss, dp, qd_lld[e], lam);
        };
        if (len == TW_CK) {
#pragma unroll 8
          for (int e = 0; e < TW_CK; ++e) step(e);
        } else {
          for (int e = 0; e < len; ++e) step(e);
        }
      } else {
        double pp = (c + 1 < nck) ? Pck[(int64_t)(c + 1) * T] : 0.0;  // p_{c0+len}
        auto step = [&](int e) {
          const double dm = guard(qd_lld[e] + pp, pivmin);
          pp = prog_prev(pp, qd_d[e], dm, qd_lld[e], lam);
          bw[e][lane] = pp;
        };
        if (len == TW_CK && c0 + TW_CK - 1 < n - 1) {
#pragma unroll 8
          for (int e = TW_CK - 1; e >= 0; --e) step(e);
        } else {
          for (int e = len - 1; e >= 0; --e) {
            if (c0 + e == n - 1) {
              pp = dn1 - lam;
              bw[e][lane] = pp;
            } else {
              step(e);
            }
          }
        }
      }
      __syncthreads();
      // wave 0 takes e < 32 of the chunk, wave 1 e >= 32
#pragma unroll 4
      for (int q = 0; q < TW_CK / 2; ++q) {
        const int e = wave * (TW_CK / 2) + q;
        const int i = c0 + e;
        if (e < len && i >= g0 && i < g1) {
          const double g = buf[0][e][lane] + buf[1][e][lane] + lam;
          if (fabs(g) < fabs(gbest) || (fabs(g) == fabs(gbest) && fabs(g) < INFINITY)) {
            gbest = g;
            r = i;
          }
        }
      }
    }
    // combine: the smaller |gamma|, the larger index on ties; with no
    // finite gamma at all the twist stays at the range's last index
    {
      xg[wave][lane] = gbest;
      xr[wave][lane] = r;
      if (wave == 0) xc[lane] = cnt;
      __syncthreads();
      const double g0v = xg[0][lane], g1v = xg[1][lane];
      const int rr0 = xr[0][lane], rr1 = xr[1][lane];
      cnt = xc[lane];
      const bool take1 = rr1 >= 0 &&
                         (rr0 < 0 || fabs(g1v) < fabs(g0v) || (fabs(g1v) == fabs(g0v) && rr1 > rr0));
      if (take1) {
        gbest = g1v;
        r = rr1;
      } else if (rr0 >= 0) {
        gbest = g0v;
        r = rr0;
      } else {
        gbest = INFINITY;
        r = g1 - 1;
      }
    }
    // pass 3: wave 0 z downward from r (needs p_{i+1}), wave 1 upward (s_i)
    double nzp = wave == 0 ? 1.0 : 0.0;
    if (active && wave == 0) out[oidx(r)] = 1.0;
    {
      double z = 1.0;
      for (int c = 0; c < nck; ++c) {
        const int cw = wave == 0 ? c : nck - 1 - c;
        stage(cw);
        const int c0 = cw * TW_CK, len = min(TW_CK, n - c0);
        if (wave == 0) {
          double pp;
          int etop;
          if (cw + 1 < nck) {
            pp = Pck[(int64_t)(cw + 1) * T];  // p_{c0+len}
            etop = len - 1;
          } else {
            pp = dn1 - lam;  // p_{n-1}
            etop = len - 2;
          }
          auto pstep = [&](int e) {
            bw[e][lane] = pp;  // p_{c0+e+1}
            const double dm = guard(qd_lld[e] + pp, pivmin);
            pp = prog_prev(pp, qd_d[e], dm, qd_lld[e], lam);
          };
          if (etop == TW_CK - 1) {
#pragma unroll 8
            for (int e = TW_CK - 1; e >= 0; --e) pstep(e);
          } else {
            for (int e = etop; e >= 0; --e) pstep(e);
          }
          // z in straight-line code (the quotient formed on every step and
          // kept where the step applies), staged in bw and stored after
#pragma unroll 8
          for (int e = 0; e < TW_CK; ++e) {
            const bool go = e <= etop && active && c0 + e >= r;
            const double zn = keep(-(qd_ld[e] / guard(qd_lld[e] + bw[e][lane], pivmin)) * z);
            z = go ? zn : z;
            nzp += go ? zn * zn : 0.0;
            bw[e][lane] = zn;
          }
          for (int e = 0; e < TW_CK; ++e) {
            const int i = c0 + e;
            if (e <= etop && active && i >= r) out[oidx(i + 1)] = bw[e][lane];
          }
        } else {
          double ss = Sck[(int64_t)cw * T];
          auto sstep = [&](int e) {
            bw[e][lane] = ss;
            const double dp = guard(qd_d[e] + ss, pivmin);
            ss = stat_next(ss, dp, qd_lld[e], lam);
          };
          if (len == TW_CK) {
#pragma unroll 8
            for (int e = 0; e < TW_CK; ++e) sstep(e);
          } else {
            for (int e = 0; e < len; ++e) sstep(e);
          }
#pragma unroll 8
          for (int e = TW_CK - 1; e >= 0; --e) {
            const bool go = e < len && active && c0 + e < r;
            const double zn = keep(-(qd_ld[e] / guard(qd_d[e] + bw[e][lane], pivmin)) * z);
            z = go ? zn : z;
            nzp += go ? zn * zn : 0.0;
            bw[e][lane] = zn;
          }
          for (int e = TW_CK - 1; e >= 0; --e) {
            const int i = c0 + e;
            if (e < len && active && i < r) out[oidx(i)] = bw[e][lane];
          }
        }
      }
    }
    xn[wave][lane] = nzp;
    __syncthreads();
    // converged lanes keep the norm of their final vector
    if (active) {
      nz2 = xn[0][lane] + xn[1][lane];
      rfin = r;
    }
    if (!active) continue;
    corr = gbest / nz2;
    // converged: a value at the zero floor (lam < 0, its first vector), a
    // correction within a few ulps of lam (the shift's error is
    // then far below any gap the vector can resolve, and the value returned
    // includes the correction), the vector's norm saturated, or (from the
    // third step on) the correction no longer halving -- it has reached the
    // rounding floor of gamma_r, which for small singular values sits above
    // eps*lam
    if (lam < 0.0 || fabs(corr) <= TW_CONV * TW_EPS * lam || nz2 > 1.0 / (TW_EPS * TW_EPS) ||
        (it >= 2 && fabs(corr) > 0.5 * prev)) {
      active = false;
      continue;
    }
    prev = fabs(corr);
    // cnt = number of eigenvalues below lam: keep the bracket around idx
    // (an inseparable cluster's bracket is already at the rounding level)
    if (!unres) {
      if (cnt <= idx) lo = lam; else hi = lam;
    }
    double nl = lam + corr;
    if (!(nl > lo && nl < hi)) nl = 0.5 * (lo + hi);
    lam = nl;
  }
  __syncthreads();  // both waves' vector entries are written
  if (!live) return;
  const double inv = 1.0 / sqrt(nz2);
  const int jh = (n + 1) / 2;
  const int ja = wave == 0 ? 0 : jh, jb = wave == 0 ? jh : n;
  for (int j0 = ja; j0 < jb; j0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = out[(int64_t)min(j0 + u, jb - 1) * ldv];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < jb) out[(int64_t)(j0 + u) * ldv] = v[u] * inv;
  }
  if (rep == 0 && wave == 0) {
    a.sig_out[idx] = sqrt(fmax(lam + corr, 0.0));
    a.blk[idx] = a.bs[min(max(rfin, 0), n - 1)];
  }
}

// ---------------------------------------------------------------------------
// Rayleigh-Ritz over groups of close values in one block
// ---------------------------------------------------------------------------

struct RrArgs {
  int n;
  const double* qd;       // both representations
  const double* a;        // the bidiagonal (split as in qd)
  const double* b;
  double tau_split;
  const int4* groups;     // per window: {member offset, k, block start, block end}
  const int4* wins;       // per window: {offset of its group's first member,
                          //  members of the group before it, group size, 0}
  int base;               // first window of this launch
  const int* members;     // global indices, ascending within a group
  double* sigma;          // refined values (in/out)
  double* Vt;
  double* Ut;
  int ldv;
  double* ck;             // checkpoints [2][n][nck]
  int nck;
  double smax;
};

// deterministic value in [-1, 1) from two integers (splitmix64 finaliser)
__device__ __forceinline__ double tw_hash(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Work-group sum of one value per thread (256 threads); every thread gets it.
__device__ __forceinline__ double wg_sum256(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  const double t = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return t;
}

// One work-group (4 waves) per group of k <= TW_KMAX members sharing the
// block [r0, r1).  Waves 0 and 1 run inverse iteration for the right (rep 0)
// and left (rep 1) vectors, one member per lane; the whole work-group then
// orthonormalises both bases, forms M = U_c^T B V_c, diagonalises it by
// one-sided Jacobi and rotates the bases.
__global__ __launch_bounds__(256) void tw_rr(RrArgs A) {
  __shared__ double qs[2][3][TW_CK];
  __shared__ double buf[2][TW_CK][64];  // recomputed pivots; later M and Vm
  __shared__ double tau_s[TW_KMAX];
  __shared__ double red[8];
  __shared__ double coef[TW_KMAX];
  __shared__ double sv[TW_KMAX];
  __shared__ int perm[TW_KMAX];
  __shared__ double part[4][64];
  __shared__ double wtmp[TW_KMAX];
  __shared__ int flag;
  const int4 g = A.groups[A.base + blockIdx.x];
  const int4 gw = A.wins[A.base + blockIdx.x];
  const int moff = g.x, k = g.y, r0 = g.z, r1 = g.w;
  const int* pmem = A.members + gw.x;  // the group's earlier windows (final)
  const int np = gw.y;
  const int n = A.n, L = r1 - r0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ldv = A.ldv;
  const int* mem = A.members + moff;
  const double smax = A.smax;
  const double pivmin = TW_TINY * fmax(1.0, smax * smax);
  // a group wider than one window takes its left basis from the right one
  // (phase B), unless its values are at the rounding level of zero
  const bool fromv = gw.z > TW_KMAX && A.sigma[mem[0]] > 1024.0 * TW_EPS * smax;

  // shifts (in sigma^2).  A member whose neighbours in the group lie more
  // than TW_TIGHT eps away takes its own value.  Members closer than that
  // cannot be told apart by their own shifts (the values are accurate to a
  // few eps): a shift within ulps of a run of such values amplifies
  // whichever vector of the run lies nearest, every member's iterate
  // collapses onto it, and Gram-Schmidt of nearly parallel vectors loses the
  // run's other directions.  The members of such a run share one shift
  // below the run, by twice its width plus 256 eps (and twice the floor,
  // for runs of zero values), so that every vector of the run is amplified
  // alike and each member's random start keeps its own mixture of them; the
  // shift stays nearer the run than half the gap to the member below.
  if (tid == 0) {
    const double floor2 = TW_EPS * TW_EPS * smax * smax;
    for (int q = 0; q < k; ++q) {
      const double s = A.sigma[mem[q]];
      wtmp[q] = fmax(s * s, floor2);
    }
    for (int q = 0; q < k;) {
      int e = q;
      while (e + 1 < k && wtmp[e + 1] - wtmp[e] <= TW_TIGHT * TW_EPS * wtmp[e + 1]) ++e;
      double sh = wtmp[q];
      if (e > q) {
        double dl = 2.0 * (wtmp[e] - wtmp[q]) + 256.0 * TW_EPS * wtmp[e] + 2.0 * floor2;
        if (q > 0) dl = fmin(dl, 0.5 * (wtmp[q] - wtmp[q - 1]));
        sh = wtmp[q] - dl;
      }
      for (int r = q; r <= e; ++r) tau_s[r] = sh;
      q = e + 1;
    }
  }
  __syncthreads();

  // ---- phase A: TW_RRIT steps of inverse iteration per member ----
  // waves 0 and 1 (rep 0 and rep 1) compute; waves 2 and 3 only keep the
  // barriers company
  {
    const int rep = wave & 1;
    const bool wk = wave < 2;
    const double* d = A.qd + (size_t)rep * 3 * n;
    const double* ld = d + n;
    const double* lld = d + 2 * (size_t)n;
    const int q0 = rep ? n - r1 : r0;  // block in this representation
    const bool act = wk && lane < k;
    const int gi = act ? mem[lane] : mem[0];
    double* vec = (rep ? A.Ut : A.Vt) + gi;
    auto vi = [&](int e) -> int64_t { return (int64_t)(rep ? (n - 1 - e) : e) * ldv; };
    double* ckp = A.ck + ((size_t)rep * n + gi) * A.nck;
    const double tau = act ? tau_s[lane] : 1.0;
    const double pm = fmax(pivmin, TW_EPS * tau);
    const int nc = (L + TW_CK - 1) / TW_CK;
    double scale = 1.0;
    for (int itr = 0; itr < TW_RRIT; ++itr) {
      // forward: pivots of L D L^T - tau I = L+ D+ L+^T, w = L+^{-1} x,
      // y' = w / D+ stored in place
      double s = -tau, wprev = 0.0, lprev = 0.0;
      for (int c = 0; c < nc; ++c) {
        const int e0 = q0 + c * TW_CK, len = min(TW_CK, L - c * TW_CK);
        __syncthreads();
        if (wk && lane < len) {
          qs[rep][0][lane] = d[e0 + lane];
          qs[rep][1][lane] = ld[e0 + lane];
          qs[rep][2][lane] = lld[e0 + lane];
        }
        __syncthreads();
        if (act) {
          ckp[c] = s;
          for (int e = 0; e < len; ++e) {
            const int i = e0 + e;
            const double x =
                itr == 0 ? tw_hash(((uint64_t)gi << 32) ^ (uint64_t)(i + rep * n))
                         : vec[vi(i)] * scale;
            const double w = x - lprev * wprev;
            const double dp = guard(qs[rep][0][e] + s, pm);
            vec[vi(i)] = w / dp;
            lprev = qs[rep][1][e] / dp;
            wprev = w;
            s = stat_next(s, dp, qs[rep][2][e], tau);
          }
        }
      }
      // backward: y_i = y'_i - (ld_i / D+_i) y_{i+1}, pivots recomputed per
      // chunk from the checkpoints
      double ynext = 0.0, nrm = 0.0;
      for (int c = nc - 1; c >= 0; --c) {
        const int e0 = q0 + c * TW_CK, len = min(TW_CK, L - c * TW_CK);
        __syncthreads();
        if (wk && lane < len) {
          qs[rep][0][lane] = d[e0 + lane];
          qs[rep][1][lane] = ld[e0 + lane];
          qs[rep][2][lane] = lld[e0 + lane];
        }
        __syncthreads();
        if (act) {
          double ss = ckp[c];
          for (int e = 0; e < len; ++e) {
            const double dp = guard(qs[rep][0][e] + ss, pm);
            buf[rep][e][lane] = qs[rep][1][e] / dp;  // L+_i
            ss = stat_next(ss, dp, qs[rep][2][e], tau);
          }
          for (int e = len - 1; e >= 0; --e) {
            const int i = e0 + e;
            double y = vec[vi(i)];
            if (i != q0 + L - 1) y -= buf[rep][e][lane] * ynext;
            vec[vi(i)] = y;
            nrm += y * y;
            ynext = y;
          }
        }
      }
      scale = (nrm > 0.0 && nrm < INFINITY) ? 1.0 / sqrt(nrm) : 1.0;
    }
    if (act)
      for (int e = 0; e < L; ++e) vec[vi(q0 + e)] *= scale;
  }
  __syncthreads();

  // ---- phase B: Gram-Schmidt, twice, on both bases ----
  // v_q -= sum_p (v_p . v_q) v_p over the members of the group's earlier
  // windows (final: earlier launches) and then the window's earlier members,
  // 64 at a time, then normalise; rows of the block are spread over the
  // threads.  A group wider than one window (fromv) takes its left basis
  // from the right one, U_c = B V_c before the orthonormalisation: in a run
  // of values that no shift separates and that is wider than the window,
  // independent left and right iterates span different subspaces of the
  // run, and U_c^T B V_c would not hold its values.
  auto member = [&](int p) { return p < np ? pmem[p] : mem[p - np]; };
  for (int rep = 0; rep < 2; ++rep) {
    double* M = rep ? A.Ut : A.Vt;
    if (rep == 1 && fromv) {
      for (int e = wave; e < L; e += 4) {
        const int i = r0 + e;
        const int64_t row = (int64_t)i * ldv;
        if (lane < k) {
          double bv = A.a[i] * A.Vt[row + mem[lane]];
          if (i + 1 < r1) {
            const double bi = A.b[i];
            bv += (fabs(bi) <= A.tau_split ? 0.0 : bi) * A.Vt[row + ldv + mem[lane]];
          }
          A.Ut[row + mem[lane]] = bv;
        }
      }
      __syncthreads();
    }
    for (int q = 0; q < k; ++q) {
      for (int pass = 0; pass < 2; ++pass) {
        // dots with earlier members: lane p of every wave, rows split over
        // the four waves
        for (int p0 = 0; p0 < np + q; p0 += 64) {
          const int cnt = min(64, np + q - p0);
          double acc = 0.0;
          if (lane < cnt) {
            const int mp = member(p0 + lane);
            for (int e = wave; e < L; e += 4) {
              const int64_t row = (int64_t)(r0 + e) * ldv;
              acc += M[row + mp] * M[row + mem[q]];
            }
          }
          __syncthreads();
          buf[0][wave][lane] = acc;
          __syncthreads();
          if (tid < cnt)
            coef[tid] = buf[0][0][tid] + buf[0][1][tid] + buf[0][2][tid] + buf[0][3][tid];
          __syncthreads();
          for (int e = tid; e < L; e += 256) {
            const int64_t row = (int64_t)(r0 + e) * ldv;
            double v = M[row + mem[q]];
            for (int p = 0; p < cnt; ++p) v -= coef[p] * M[row + member(p0 + p)];
            M[row + mem[q]] = v;
          }
          __syncthreads();
        }
        double nn = 0.0;
        for (int e = tid; e < L; e += 256) {
          const double v = M[(int64_t)(r0 + e) * ldv + mem[q]];
          nn += v * v;
        }
        nn = wg_sum256(nn, red);
        const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        for (int e = tid; e < L; e += 256) M[(int64_t)(r0 + e) * ldv + mem[q]] *= inv;
        __syncthreads();
      }
    }
  }

  // ---- phase C: M = U_c^T B V_c (k x k), column-major in buf ----
  double* Mm = &buf[0][0][0];             // Mm[q * TW_KMAX + p] = M[p][q]
  double* Vm = &buf[1][0][0];             // Vm[q * TW_KMAX + p]
  for (int q = 0; q < k; ++q) {
    // lane p of every wave: rows of the block split over the four waves
    double acc = 0.0;
    if (lane < k)
      for (int e = wave; e < L; e += 4) {
        const int i = r0 + e;
        const int64_t row = (int64_t)i * ldv;
        double bv = A.a[i] * A.Vt[row + mem[q]];
        if (i + 1 < r1) {
          const double bi = A.b[i];
          bv += (fabs(bi) <= A.tau_split ? 0.0 : bi) * A.Vt[row + ldv + mem[q]];
        }
        acc += A.Ut[row + mem[lane]] * bv;
      }
    part[wave][lane] = acc;
    __syncthreads();
    if (tid < k) Mm[q * TW_KMAX + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
  }
  for (int e = tid; e < TW_KMAX * TW_KMAX; e += 256) {
    const int q = e / TW_KMAX, p = e % TW_KMAX;
    Vm[e] = (p == q) ? 1.0 : 0.0;
  }
  __syncthreads();

  // ---- phase D: one-sided Jacobi on the columns of M ----
  // round-robin pairing of kk = k rounded up to even columns; lane l of
  // wave 0 takes pair l of a round; a column index >= k is a dummy
  const int kk = (k + 1) & ~1;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) flag = 0;
    __syncthreads();
    for (int rd = 0; rd < kk - 1; ++rd) {
      if (wave == 0 && lane < kk / 2) {
        // circle method: position 0 fixed, others rotate
        auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + rd) % (kk - 1); };
        int p = at(lane), q = at(kk - 1 - lane);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q < k) {
          double al = 0.0, be = 0.0, ga = 0.0;
          for (int r = 0; r < k; ++r) {
            const double x = Mm[p * TW_KMAX + r], y = Mm[q * TW_KMAX + r];
            al += x * x;
            be += y * y;
            ga += x * y;
          }
          if (fabs(ga) > TW_EPS * sqrt(al * be) && fabs(ga) > TW_TINY) {
            flag = 1;
            const double zeta = (be - al) / (2.0 * ga);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
            for (int r = 0; r < k; ++r) {
              const double x = Mm[p * TW_KMAX + r], y = Mm[q * TW_KMAX + r];
              Mm[p * TW_KMAX + r] = cs * x - sn * y;
              Mm[q * TW_KMAX + r] = sn * x + cs * y;
              const double u = Vm[p * TW_KMAX + r], w = Vm[q * TW_KMAX + r];
              Vm[p * TW_KMAX + r] = cs * u - sn * w;
              Vm[q * TW_KMAX + r] = sn * u + cs * w;
            }
          }
        }
      }
      __syncthreads();
    }
    if (flag == 0) break;
    __syncthreads();
  }

  // ---- phase E: singular values, ascending order, left rotation ----
  if (tid < k) {
    double s2 = 0.0;
    for (int r = 0; r < k; ++r) s2 += Mm[tid * TW_KMAX + r] * Mm[tid * TW_KMAX + r];
    sv[tid] = sqrt(s2);
  }
  __syncthreads();
  if (tid < k) {
    int rank = 0;
    for (int p = 0; p < k; ++p) rank += (sv[p] < sv[tid]) || (sv[p] == sv[tid] && p < tid);
    perm[rank] = tid;  // perm[position] = Jacobi column
  }
  __syncthreads();
  // Um: columns of M Vm normalised, made orthonormal by Gram-Schmidt in
  // descending order of the singular values (columns of a zero singular
  // value are completed with unit vectors); kept in Mm
  if (tid == 0) {
    for (int pos = k - 1; pos >= 0; --pos) {
      const int q = perm[pos];
      double* u = &Mm[q * TW_KMAX];
      double best = -1.0;
      for (int tryc = -1; tryc < k; ++tryc) {
        double* w = wtmp;
        for (int r = 0; r < k; ++r) w[r] = tryc < 0 ? u[r] : (r == tryc ? 1.0 : 0.0);
        for (int rp = 0; rp < 2; ++rp)
          for (int pp = k - 1; pp > pos; --pp) {
            const double* o = &Mm[perm[pp] * TW_KMAX];
            double c = 0.0;
            for (int r = 0; r < k; ++r) c += o[r] * w[r];
            for (int r = 0; r < k; ++r) w[r] -= c * o[r];
          }
        double nw = 0.0;
        for (int r = 0; r < k; ++r) nw += w[r] * w[r];
        nw = sqrt(nw);
        const double thr = tryc < 0 ? 1e-3 * sv[q] : 0.0;
        if (tryc < 0 && nw > thr && sv[q] > 64.0 * TW_EPS * sv[perm[k - 1]]) {
          for (int r = 0; r < k; ++r) u[r] = w[r] / nw;
          best = 1.0;
          break;
        }
        if (tryc >= 0 && nw > best) {
          best = nw;
          for (int r = 0; r < k; ++r) coef[r] = w[r] / nw;
        }
        if (tryc == k - 1)
          for (int r = 0; r < k; ++r) u[r] = coef[r];
      }
    }
  }
  __syncthreads();
  // rotate: V_c <- V_c Vm(:, perm), U_c <- U_c Um(:, perm); one row of the
  // block per wave at a time, lane = output column
  for (int rep = 0; rep < 2; ++rep) {
    double* X = rep ? A.Ut : A.Vt;
    const double* R = rep ? Mm : Vm;
    for (int e = wave; e < L; e += 4) {
      const int64_t row = (int64_t)(r0 + e) * ldv;
      double acc = 0.0;
      if (lane < k) {
        const int qc = perm[lane];
        for (int p = 0; p < k; ++p) acc += X[row + mem[p]] * R[qc * TW_KMAX + p];
      }
      // every lane of the wave has read the row before any lane writes it
      __builtin_amdgcn_wave_barrier();
      if (lane < k) X[row + mem[lane]] = acc;
    }
  }
  if (tid < k) A.sigma[mem[tid]] = sv[perm[tid]];
}

// u_i^T B v_i >= 0.  Lane = singular value i; the four waves of a group
// split the index range j and their partial sums meet in LDS.  Loads are
// issued eight indices at a time so that their latencies overlap.
__global__ __launch_bounds__(256) void tw_sign(int n, const double* a, const double* b,
                                               const double* Vt, double* Ut, int ldv) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int ic = i < n ? i : n - 1;  // a valid column for the loads of idle lanes
  const int q = (n + 3) / 4;
  const int j0 = wave * q, j1 = min(n, j0 + q);
  double t = 0.0;
  for (int j = j0; j < j1; j += 8) {
    double av[8], bv[8], vj[8], vn[8], uj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int jj = min(j + u, n - 1);         // clamped: in bounds, unused past j1
      const int jn = min(jj + 1, n - 1);
      av[u] = a[jj];
      bv[u] = jj < n - 1 ? b[jj] : 0.0;
      vj[u] = Vt[(int64_t)jj * ldv + ic];
      vn[u] = Vt[(int64_t)jn * ldv + ic];
      uj[u] = Ut[(int64_t)jj * ldv + ic];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j + u < j1) t += uj[u] * (av[u] * vj[u] + bv[u] * vn[u]);
  }
  part[wave][lane] = t;
  __syncthreads();
  const double tot = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
  if (i < n && tot < 0.0)
    for (int j = j0; j < j1; ++j) Ut[(int64_t)j * ldv + i] = -Ut[(int64_t)j * ldv + i];
}

// sigma[i] = max(sigma[0..i]): values of an inseparable cluster may come out
// of the refinement a rounding error out of order.
__global__ __launch_bounds__(1024) void tw_mono(int n, double* sigma) {
  __shared__ double sh[1024];
  __shared__ double carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0.0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + t;
    double v = i < n ? sigma[i] : 0.0;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const double w = t >= o ? sh[t - o] : v;
      __syncthreads();
      v = fmax(v, w);
      sh[t] = v;
      __syncthreads();
    }
    v = fmax(v, carry);
    if (i < n) sigma[i] = v;
    __syncthreads();
    if (t == 1023) carry = v;
    __syncthreads();
  }
}

// out[0] = max |a_i|, |b_i| (i < n-1); out[1] = sigma[n-1]; out[2] = 1 if
// a value is not finite.  One work-group of 1024 threads.
__global__ __launch_bounds__(1024) void tw_scale(int n, const double* a, const double* b,
                                                 const double* sigma, double* out) {
  __shared__ double sh[1024];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  double m = 0.0;
  for (int i = t; i < n; i += 1024) {
    m = fmax(m, fabs(a[i]));
    if (i < n - 1) m = fmax(m, fabs(b[i]));
    if (!isfinite(sigma[i])) bad = 1;
  }
  sh[t] = m;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) sh[t] = fmax(sh[t], sh[t + o]);
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sh[0];
    out[1] = sigma[n - 1];
    out[2] = bad;
  }
}

// B = 0: sigma = 0, U = V = I.
__global__ void tw_identity(int n, double* Vt, double* Ut, int ldv, double* sigma) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int j = (int)(e / n), i = (int)(e % n);
  Vt[(int64_t)j * ldv + i] = Ut[(int64_t)j * ldv + i] = (i == j) ? 1.0 : 0.0;
  if (j == 0) sigma[i] = 0.0;
}

static size_t tw_ck_size(int n) { return (size_t)((n + TW_CK - 1) / TW_CK) * 2 * n; }

void twisted_plan(ArenaPlan& p, int n) {
  p.add<double>(6 * (size_t)n);        // qd
  p.add<double>(2 * tw_ck_size(n));    // checkpoints
  p.add<double>(n);                    // D&C values
  p.add<double>(2 * (size_t)n);        // brackets
  p.add<int>(9 * (size_t)n + 64);      // split, bs, be, cs, ce, r0, r1, flags, blk, count
  p.add<int>(4 * (size_t)n + 4);       // window table (int4 per window, at most n / 2 + n / 64)
  p.add<int>(4 * (size_t)n + 4);       // window's group (int4 per window)
  p.add<int>(n);                       // group members
}

int twisted_run(int n, const double* a, const double* b, double* sigma,
                double* Vt, double* Ut, int ldv, Arena& ws, hipStream_t st) {
  if (n < 1 || ldv < n) return DDC_E_ARG;
  double* qd = ws.take<double>(6 * (size_t)n);
  double* scr = ws.take<double>(2 * tw_ck_size(n));
  double* sig0 = ws.take<double>(n);
  double* lam = ws.take<double>(2 * (size_t)n);
  int* ib = ws.take<int>(9 * (size_t)n + 64);
  int4* gtab = ws.take<int4>((size_t)n + 1);
  int4* gwin = ws.take<int4>((size_t)n + 1);
  int* gmem = ws.take<int>(n);
  if (!gmem) return DDC_E_NOMEM;
  int* split = ib;
  int* bs = ib + n;
  int* be = ib + 2 * (size_t)n;
  int* cs = ib + 3 * (size_t)n;
  int* ce = ib + 4 * (size_t)n;
  int* r0 = ib + 5 * (size_t)n;
  int* r1 = ib + 6 * (size_t)n;
  int* flags = ib + 7 * (size_t)n;
  int* blk = ib + 8 * (size_t)n;
  int* cnt = ib + 9 * (size_t)n;
  DDC_HIP(hipMemcpyAsync(sig0, sigma, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(tw_scale, dim3(1), dim3(1024), 0, st, n, a, b, sigma, scr);
  DDC_LAUNCHED();
  double hsc[3] = {0.0, 0.0, 0.0};
  DDC_HIP(hipMemcpyAsync(hsc, scr, sizeof(hsc), hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  if (hsc[2] != 0.0 || !(hsc[0] < INFINITY)) {
    report_error("singular values of the bidiagonal are not finite");
    return DDC_E_NUMERIC;
  }
  double smax = hsc[1];
  if (hsc[0] == 0.0) {
    // B = 0
    hipLaunchKernelGGL(tw_identity, dim3((unsigned)ceil_div((int64_t)n * n, 256)), dim3(256), 0,
                       st, n, Vt, Ut, ldv, sigma);
    DDC_LAUNCHED();
    return DDC_OK;
  }
  if (!(smax > 0.0)) smax = hsc[0];  // a bound of sigma_max within a factor 2
  const double tau_split = TW_SPLIT * TW_EPS * smax;
  const double dl = 8.0 * n * TW_EPS * smax;
  hipLaunchKernelGGL(tw_prep, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n, a, b,
                     tau_split, qd, split);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(tw_scan, dim3(1), dim3(1024), 0, st, n, split, sig0, dl, bs, be, cs, ce, cnt);
  DDC_LAUNCHED();
  IsoArgs ia;
  ia.n = n;
  ia.qd = qd;
  ia.sig = sig0;
  ia.bs = bs;
  ia.be = be;
  ia.cs = cs;
  ia.ce = ce;
  ia.cnt = cnt;
  ia.smax = smax;
  ia.dl = dl;
  ia.lam_lo = lam;
  ia.lam_hi = lam + n;
  ia.r0 = r0;
  ia.r1 = r1;
  ia.flags = flags;
  hipLaunchKernelGGL(tw_isolate, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, ia);
  DDC_LAUNCHED();
  TwArgs t;
  t.n = n;
  t.qd = qd;
  t.sig_in = sig0;
  t.sig_out = sigma;
  t.Vt = Vt;
  t.Ut = Ut;
  t.ldv = ldv;
  t.T = 2 * n;
  t.Sck = scr;
  t.Pck = scr + tw_ck_size(n);
  t.smax = smax;
  t.lam_lo = lam;
  t.lam_hi = lam + n;
  t.r0 = r0;
  t.r1 = r1;
  t.flags = flags;
  t.bs = bs;
  t.blk = blk;
  hipLaunchKernelGGL(tw_solve, dim3(2 * (unsigned)ceil_div(n, 64)), dim3(128), 0, st, t);
  DDC_LAUNCHED();

  // groups: lanes of one block whose refined values are within TW_GTOL
  // (relative, in sigma^2) of the block's previous lane
  std::vector<int> hblk(n), hbe(n);
  std::vector<double> hs(n);
  DDC_HIP(hipMemcpyAsync(hblk.data(), blk, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipMemcpyAsync(hbe.data(), be, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipMemcpyAsync(hs.data(), sigma, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  // per block: its last lane, the first lane and size of its open group;
  // members of a group are chained through nxt
  std::vector<int> last(n, -1), gst(n, -1), gsz(n, 0), nxt(n, -1);
  // A group is cut into windows of TW_KMAX members (the last may hold
  // one); window w of a group runs in launch w, after the group's earlier
  // windows, against whose final vectors it is orthogonalised.
  std::vector<int4> tab, win;
  std::vector<int> mem, cur, lvl;
  auto close = [&](int kb) {
    if (gsz[kb] < 2) return;
    cur.clear();
    for (int j = gst[kb]; j >= 0; j = nxt[j]) cur.push_back(j);
    const int g0 = (int)mem.size();
    mem.insert(mem.end(), cur.begin(), cur.end());
    for (size_t w = 0; w < cur.size(); w += TW_KMAX) {
      const int k = (int)std::min<size_t>(TW_KMAX, cur.size() - w);
      tab.push_back(make_int4(g0 + (int)w, k, kb, hbe[kb]));
      win.push_back(make_int4(g0, (int)w, (int)cur.size(), 0));
      lvl.push_back((int)(w / TW_KMAX));
    }
  };
  for (int j = 0; j < n; ++j) {
    const int kb = std::min(std::max(hblk[j], 0), n - 1);
    const int p = last[kb];
    const double s2 = hs[j] * hs[j];
    if (p >= 0 && s2 - hs[p] * hs[p] <= TW_GTOL * s2) {
      nxt[p] = j;
      ++gsz[kb];
    } else {
      if (p >= 0) close(kb);
      gst[kb] = j;
      gsz[kb] = 1;
    }
    last[kb] = j;
  }
  for (int kb = 0; kb < n; ++kb)
    if (last[kb] >= 0) close(kb);
  if (!tab.empty()) {
    // windows in launch order: every group's first window, then the second ...
    int nlev = 0;
    for (int l : lvl) nlev = std::max(nlev, l + 1);
    std::vector<int4> tabl, winl;
    std::vector<int> lstart(nlev + 1, 0);
    for (int l = 0; l < nlev; ++l) {
      lstart[l] = (int)tabl.size();
      for (size_t i = 0; i < tab.size(); ++i)
        if (lvl[i] == l) {
          tabl.push_back(tab[i]);
          winl.push_back(win[i]);
        }
    }
    lstart[nlev] = (int)tabl.size();
    DDC_HIP(hipMemcpyAsync(gtab, tabl.data(), sizeof(int4) * tabl.size(), hipMemcpyHostToDevice, st));
    DDC_HIP(hipMemcpyAsync(gwin, winl.data(), sizeof(int4) * winl.size(), hipMemcpyHostToDevice, st));
    DDC_HIP(hipMemcpyAsync(gmem, mem.data(), sizeof(int) * mem.size(), hipMemcpyHostToDevice, st));
    RrArgs ra;
    ra.n = n;
    ra.qd = qd;
    ra.a = a;
    ra.b = b;
    ra.tau_split = tau_split;
    ra.groups = gtab;
    ra.wins = gwin;
    ra.members = gmem;
    ra.sigma = sigma;
    ra.Vt = Vt;
    ra.Ut = Ut;
    ra.ldv = ldv;
    ra.ck = scr;
    ra.nck = (n + TW_CK - 1) / TW_CK;
    ra.smax = smax;
    for (int l = 0; l < nlev; ++l) {
      ra.base = lstart[l];
      hipLaunchKernelGGL(tw_rr, dim3((unsigned)(lstart[l + 1] - lstart[l])), dim3(256), 0, st, ra);
      DDC_LAUNCHED();
    }
    // the host vectors must outlive the copies
    DDC_HIP(hipStreamSynchronize(st));
  }
  hipLaunchKernelGGL(tw_sign, dim3((unsigned)ceil_div(n, 64)), dim3(256), 0, st, n, a, b, Vt,
                     Ut, ldv);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(tw_mono, dim3(1), dim3(1024), 0, st, n, sigma);
  DDC_LAUNCHED();
  return DDC_OK;
}

}  // namespace ddc
