// dc.hip -- singular values of an upper bidiagonal by divide and conquer, FP64.
//
// Replaces GetSingularValues_Parallel (Calculations-Parallel.c:852-874) and
// its recursion (DivideAndConquer_Parallel :706-850).  The split is the
// reference's: B is N x (N+1) with diagonal a and super-diagonal b, row
// K = N/2 is removed, the children are K x (K+1) and (N-K-1) x (N-K), and
// each merge needs only the first and last rows of the children's right
// singular vector matrices (the Konda-Nakamura values-only form).
//
// What differs from the reference is the numerical core of a merge:
//  * relative deflation (|z_j| or pole gaps below 8 eps scale) with Givens
//    rotations of the first/last rows, instead of the absolute thresholds
//    1e-14/1e-20 (:143-144, :467);
//  * each secular root is bracketed between its poles and found by bisection
//    in the offset tau from the nearer pole (geometric while the bracket
//    spans orders of magnitude), so sigma - d_j is formed without
//    cancellation; no Newton steps that can jump to a neighbouring root
//    (the reference's duplicate-root failure, :163-254);
//  * the first/last rows come from the Gu-Eisenstat z-hat (:445-465 uses a
//    log-domain product with the same role), evaluated from the stored
//    offsets.
//
// GPU mapping: the recursion tree is planned on the host and processed
// bottom-up, one level per round.  Per level five kernels run over all
// merges of the level: prep (one work-group per merge: rank-merge of the two
// sorted child spectra, deflation, compaction), roots (one wave per root),
// zhat (one wave per pole), rows (one wave per root: first/last rows of the
// new singular vectors) and assemble (one work-group per merge).  The roots
// kernel gives a whole wave to each root: its lanes split the O(K) secular
// sum, so the top merge of an n = 8192 problem fills the GPU.
#include <algorithm>
#include <vector>

#include "ddc_common.h"

namespace ddc {

constexpr double EPS = 2.220446049250313e-16;
constexpr double TINY = 2.2250738585072014e-308;

struct DcNode {
  int off, N, K;       // positions [off, off+N), split row off+K
  int c1, c2;          // child node ids (-1 for leaves)
  int height;
};

// Per-merge device descriptor.
struct DcMerge {
  int off, N, K;
  int node, c1, c2;
  int top;  // 1 for the root merge (values only)
};

// Per-merge scalars produced by prep.
struct DcMergeState {
  int kept;          // K': non-deflated poles, including the pole at 0
  double scale;      // poles and z are divided by it in the secular solve
  double znorm;      // |z| (scaled)
  double cg, sg;     // rotation of the two null columns
};

struct DcBufs {
  double *a, *b;                 // inputs
  double *sig, *fr, *lr;         // node outputs at node positions
  double *phi, *psi;             // per node
  double *wd, *wz, *wF, *wL;     // sorted merged data
  int* flag;                     // 1 = deflated
  double *kd, *kz, *kF, *kL;     // kept (scaled poles/z)
  double *dd, *dF, *dL;          // deflated (unscaled poles)
  int* org;                      // root origin (index into kept)
  double* tau;                   // root offset from its origin pole (scaled)
  double* zh;                    // z-hat
  double *fn, *ln;               // new first/last rows for the roots
  DcMergeState* ms;              // per merge
  const DcMerge* merges;         // all merges, level-major
  const int2* items;             // (merge index, first item) per work-group
  double* out;                   // final singular values
};

// ---------------------------------------------------------------------------
// leaves: N == 1 ([a b], sigma = hypot) and N == 0 (W = [1])
// ---------------------------------------------------------------------------
__global__ void dc_leaves(DcBufs B, const int3* leaves, int nleaf) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nleaf) return;
  int off = leaves[t].x, N = leaves[t].y, node = leaves[t].z;
  if (N == 0) {
    B.phi[node] = 1.0;
    B.psi[node] = 1.0;
    return;
  }
  double a = B.a[off], b = B.b[off];
  double s = hypot(a, b);
  B.sig[off] = s;
  if (s == 0.0) {
    B.fr[off] = 1.0;
    B.lr[off] = 0.0;
    B.phi[node] = 0.0;
    B.psi[node] = 1.0;
  } else {
    // right singular vector (a, b)/s, null vector (-b, a)/s
    B.fr[off] = a / s;
    B.lr[off] = b / s;
    B.phi[node] = -b / s;
    B.psi[node] = a / s;
  }
}

// ---------------------------------------------------------------------------
// helpers for one work-group of 256 threads
// ---------------------------------------------------------------------------
__device__ __forceinline__ int count_less(const double* x, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int count_leq(const double* x, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Exclusive scan of per-thread counts over the 256-thread group.
__device__ int group_exscan(int v, int* tmp, int* total) {
  int tid = threadIdx.x;
  tmp[tid] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    int add = (tid >= o) ? tmp[tid - o] : 0;
    __syncthreads();
    tmp[tid] += add;
    __syncthreads();
  }
  int incl = tmp[tid];
  *total = tmp[255];
  __syncthreads();
  return incl - v;
}

__device__ double group_max(double v, double* red) {
  v = fmax(v, __shfl_xor(v, 32, 64));
  v = fmax(v, __shfl_xor(v, 16, 64));
  v = fmax(v, __shfl_xor(v, 8, 64));
  v = fmax(v, __shfl_xor(v, 4, 64));
  v = fmax(v, __shfl_xor(v, 2, 64));
  v = fmax(v, __shfl_xor(v, 1, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// prep: build the merged arrow, sort, deflate, compact.  One group per merge.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dc_prep(DcBufs B, int mbase) {
  __shared__ double red[4];
  __shared__ int tmp[256];
  __shared__ double s_tol, s_scale;
  __shared__ int s_total;
  const DcMerge M = B.merges[mbase + blockIdx.x];
  const int tid = threadIdx.x;
  const int off = M.off, N = M.N, K = M.K, N2 = N - K - 1;
  const double* s1 = B.sig + off;
  const double* f1 = B.fr + off;
  const double* l1 = B.lr + off;
  const double* s2 = B.sig + off + K + 1;
  const double* f2 = B.fr + off + K + 1;
  const double* l2 = B.lr + off + K + 1;
  const double aK = B.a[off + K], bK = B.b[off + K];
  const double ph1 = B.phi[M.c1], ps1 = B.psi[M.c1];
  const double ph2 = B.phi[M.c2], ps2 = B.psi[M.c2];
  const double x = aK * ps1, y = bK * ph2;
  const double r0 = hypot(x, y);
  const double cg = (r0 > 0.0) ? x / r0 : 1.0;
  const double sg = (r0 > 0.0) ? y / r0 : 0.0;

  double* wd = B.wd + off;
  double* wz = B.wz + off;
  double* wF = B.wF + off;
  double* wL = B.wL + off;
  int* flag = B.flag + off;

  // 1. merge the two ascending child spectra behind the pole at 0 (on ties
  // the first child's value first).  Merge path: thread t writes the merged
  // outputs [t per, (t+1) per), finding its start in the two lists by one
  // binary search and then walking both (a binary search per element was a
  // chain of 13 dependent loads per element, most of this kernel's time).
  double zmax = 0.0;
  if (tid == 0) {
    wd[0] = 0.0;
    wz[0] = r0;
    wF[0] = cg * ph1;
    wL[0] = sg * ps2;
    zmax = fabs(r0);
  }
  {
    const int NO = K + N2;  // merged outputs after the pole
    const int per = (NO + 255) / 256;
    const int d0 = min(NO, tid * per), d1 = min(NO, d0 + per);
    int lo = max(0, d0 - N2), hi = min(d0, K);
    while (lo < hi) {  // i = the number of first-child values among outputs < d0
      const int mid = (lo + hi) >> 1;
      if (s1[mid] <= s2[d0 - mid - 1]) lo = mid + 1;
      else hi = mid;
    }
    int i = lo, j = d0 - lo;
    for (int p = d0; p < d1; ++p) {
      const bool first = j >= N2 || (i < K && s1[i] <= s2[j]);
      if (first) {
        const double zv = aK * l1[i];
        wd[1 + p] = s1[i];
        wz[1 + p] = zv;
        wF[1 + p] = f1[i];
        wL[1 + p] = 0.0;
        zmax = fmax(zmax, fabs(zv));
        ++i;
      } else {
        const double zv = bK * f2[j];
        wd[1 + p] = s2[j];
        wz[1 + p] = zv;
        wF[1 + p] = 0.0;
        wL[1 + p] = l2[j];
        zmax = fmax(zmax, fabs(zv));
        ++j;
      }
    }
  }
  zmax = group_max(zmax, red);
  __syncthreads();
  if (tid == 0) {
    double dmax = (N > 1) ? wd[N - 1] : 0.0;
    double sc = fmax(dmax, zmax);
    s_scale = sc;
    s_tol = fmax(8.0 * EPS * sc, TINY);
  }
  __syncthreads();
  const double tol = s_tol;

  // 2. deflate negligible z_j (j >= 1)
  for (int j = tid; j < N; j += 256) flag[j] = (j >= 1 && fabs(wz[j]) <= tol) ? 1 : 0;
  __syncthreads();

  // 3. close poles: sequential rotations along the sorted kept entries.
  // Chains of close poles are rare; one thread walks the list, and only if
  // some kept entry lies within tol of its predecessor in the sorted list
  // (else no pair of kept entries is that close and the walk rotates
  // nothing).
  int close = 0;
  for (int j = 1 + tid; j < N; j += 256) close |= (!flag[j] && wd[j] - wd[j - 1] <= tol) ? 1 : 0;
  close = __syncthreads_or(close);
  if (tid == 0 && close) {
    int prev = 0;
    for (int j = 1; j < N; ++j) {
      if (flag[j]) continue;
      double dj = wd[j];
      if (dj - wd[prev] <= tol) {
        double zp = wz[prev], zj = wz[j];
        double r = hypot(zp, zj);
        if (prev == 0) {
          // pole ~0: fold z_j into z_0, deflate j (its value is ~0)
          double c = zp / r, s = zj / r;
          double F0 = wF[0], Fj = wF[j], L0 = wL[0], Lj = wL[j];
          wF[0] = c * F0 + s * Fj;
          wF[j] = -s * F0 + c * Fj;
          wL[0] = c * L0 + s * Lj;
          wL[j] = -s * L0 + c * Lj;
          wz[0] = r;
          wz[j] = 0.0;
          flag[j] = 1;
          continue;
        }
        double c = zj / r, s = zp / r;
        double Fp = wF[prev], Fj = wF[j], Lp = wL[prev], Lj = wL[j];
        wF[prev] = c * Fp - s * Fj;
        wF[j] = s * Fp + c * Fj;
        wL[prev] = c * Lp - s * Lj;
        wL[j] = s * Lp + c * Lj;
        wz[prev] = 0.0;
        wz[j] = r;
        flag[prev] = 1;
      }
      prev = j;
    }
  }
  // a zero arrow (every pole and z at zero) keeps its root at zero
  if (tid == 0 && fabs(wz[0]) <= tol) wz[0] = s_scale > 0.0 ? tol : 0.0;
  __syncthreads();

  // 4. compaction: kept entries to [0, K'), deflated to [0, N-K') in order
  const int per = (N + 255) / 256;
  const int b0 = min(N, tid * per), b1 = min(N, b0 + per);
  int nk = 0;
  for (int j = b0; j < b1; ++j) nk += flag[j] ? 0 : 1;
  int kept_tot = 0;
  int kpos = group_exscan(nk, tmp, &kept_tot);
  if (tid == 0) s_total = kept_tot;
  __syncthreads();
  const int kept = s_total;
  const double sc = s_scale;
  const double isc = (sc > 0.0) ? 1.0 / sc : 1.0;
  int dpos = b0 - kpos;
  for (int j = b0; j < b1; ++j) {
    if (!flag[j]) {
      B.kd[off + kpos] = wd[j] * isc;
      B.kz[off + kpos] = wz[j] * isc;
      B.kF[off + kpos] = wF[j];
      B.kL[off + kpos] = wL[j];
      ++kpos;
    } else {
      B.dd[off + dpos] = wd[j];
      B.dF[off + dpos] = wF[j];
      B.dL[off + dpos] = wL[j];
      ++dpos;
    }
  }
  // |z| of the kept (scaled) entries
  double zz = 0.0;
  __syncthreads();
  for (int q = tid; q < kept; q += 256) {
    double v = B.kz[off + q];
    zz += v * v;
  }
  zz = wave_sum(zz);
  if ((tid & 63) == 0) red[tid >> 6] = zz;
  __syncthreads();
  if (tid == 0) {
    DcMergeState st;
    st.kept = kept;
    st.scale = (sc > 0.0) ? sc : 1.0;
    st.znorm = sqrt(red[0] + red[1] + red[2] + red[3]);
    st.cg = cg;
    st.sg = sg;
    B.ms[mbase + blockIdx.x] = st;
    B.phi[M.node] = -sg * ph1;
    B.psi[M.node] = cg * ps2;
  }
}

// Wave-cooperative secular function 1 + sum z_j^2 / ((d_j - w)(d_j + w)) at
// w = d_o + tau: lane l takes j = l, l+64, ...; fixed-order wave sum, so every
// lane gets the same value.
__device__ __forceinline__ double secular_wave(const double* d, const double* z,
                                               int K, int o, double tau, int lane) {
  const double dO = d[o];
  double f = 0.0;
  for (int j = lane; j < K; j += 64) {
    double zj = z[j];
    double dj = d[j];
    f += (zj * zj) / (((dj - dO) - tau) * (dj + dO + tau));
  }
  return 1.0 + wave_sum(f);
}

// The same with its derivative in tau: with g_j = (d_j - w)(d_j + w),
// df/dtau = 2 w sum z_j^2 / g_j^2.
__device__ __forceinline__ double2 secular_wave_d(const double* d, const double* z, int K, int o,
                                                  double tau, int lane) {
  const double dO = d[o];
  double f = 0.0, fp = 0.0;
  for (int j = lane; j < K; j += 64) {
    const double zj = z[j];
    const double dj = d[j];
    const double g = ((dj - dO) - tau) * (dj + dO + tau);
    const double t = (zj * zj) / g;
    f += t;
    fp += t / g;
  }
  return make_double2(1.0 + wave_sum(f), 2.0 * (dO + tau) * wave_sum(fp));
}

__device__ __forceinline__ double wave_prod(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v *= __shfl_xor(v, o, 64);
  return v;
}

// Work items: each group of 256 threads takes 4 consecutive items of one
// merge, one wave per item.
#define DC_ITEM()                                         \
  const int2 it = B.items[ibase + blockIdx.x];            \
  const int mi = it.x;                                    \
  const DcMerge M = B.merges[mi];                         \
  const int lane = threadIdx.x & 63;                      \
  const int item = it.y + (threadIdx.x >> 6);

// ---------------------------------------------------------------------------
// roots: one wave per secular root
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dc_roots(DcBufs B, int ibase) {
  DC_ITEM();
  const DcMergeState st = B.ms[mi];
  const int i = item;
  const int K = st.kept;
  if (i >= K) return;
  const double* d = B.kd + M.off;
  const double* z = B.kz + M.off;
  if (K == 1) {
    if (lane == 0) {
      B.org[M.off] = 0;
      B.tau[M.off] = fabs(z[0]);
    }
    return;
  }
  int o;
  double A, sgn;
  if (i < K - 1) {
    double mid = 0.5 * (d[i] + d[i + 1]);
    double fm = secular_wave(d, z, K, i, mid - d[i], lane);
    if (fm >= 0.0) {
      o = i;
      A = mid - d[i];
      sgn = 1.0;
    } else {
      o = i + 1;
      A = d[i + 1] - mid;
      sgn = -1.0;
    }
  } else {
    o = i;
    A = st.znorm;
    sgn = 1.0;
  }
  // Safeguarded Newton on |tau| = u in the bracket (lo, hi]: f is increasing
  // in tau, and every evaluation narrows the bracket by its sign.  A Newton
  // iterate inside the bracket is the next point; otherwise the bracket is
  // bisected (geometrically while it spans a factor above four, as the
  // roots lie at any relative offset from the pole).  Stops when the Newton
  // step or the bracket is at the rounding level of u.
  double lo = 0.0, hi = A, u = -1.0, res = -1.0;
  for (int itn = 0; itn < 400; ++itn) {
    double t;
    if (u > lo && u < hi) t = u;
    else if (lo == 0.0) t = hi * 0.0625;
    else if (hi > 4.0 * lo) t = sqrt(lo * hi);
    else t = 0.5 * (lo + hi);
    if (!(t > lo && t < hi)) break;
    const double2 f = secular_wave_d(d, z, K, o, sgn * t, lane);
    if (f.x == 0.0) {
      res = t;
      break;
    }
    // origin below the root: f < 0 -> root further out; origin above: f > 0
    const bool further = (sgn > 0.0) ? (f.x < 0.0) : (f.x > 0.0);
    if (further) lo = t; else hi = t;
    if (hi - lo <= 2.0 * EPS * lo) break;
    const double du = sgn * f.y;  // df/du
    u = (du != 0.0 && fabs(du) < INFINITY) ? t - f.x / du : -1.0;
    if (fabs(u - t) <= 2.0 * EPS * t && u > lo && u < hi) {
      res = u;
      break;
    }
  }
  if (lane == 0) {
    B.org[M.off + i] = o;
    B.tau[M.off + i] = sgn * (res > 0.0 ? res : 0.5 * (lo + hi));
  }
}

// ---------------------------------------------------------------------------
// zhat: one wave per kept pole.  zhat_j^2 = (w_{K-1}^2 - d_j^2)
//   * prod_{k<j} (w_k^2 - d_j^2)/(d_k^2 - d_j^2)
//   * prod_{j<=k<K-1} (w_k^2 - d_j^2)/(d_{k+1}^2 - d_j^2)
// with w_k - d_j = (d_{org(k)} - d_j) + tau_k.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dc_zhat(DcBufs B, int ibase) {
  DC_ITEM();
  const int K = B.ms[mi].kept;
  const int j = item;
  if (j >= K) return;
  const double* d = B.kd + M.off;
  const int* org = B.org + M.off;
  const double* tau = B.tau + M.off;
  const double dj = d[j];
  double prod = 1.0;
  for (int k = lane; k < K - 1; k += 64) {
    double wm = (d[org[k]] - dj) + tau[k];  // w_k - d_j
    double wp = d[org[k]] + tau[k] + dj;    // w_k + d_j
    double dk = (k < j) ? d[k] : d[k + 1];
    prod *= (wm * wp) / ((dk - dj) * (dk + dj));
  }
  prod = wave_prod(prod);
  if (lane == 0) {
    double wm = (d[org[K - 1]] - dj) + tau[K - 1];
    double wp = d[org[K - 1]] + tau[K - 1] + dj;
    prod *= wm * wp;
    B.zh[M.off + j] = copysign(sqrt(fabs(prod)), B.kz[M.off + j]);
  }
}

// ---------------------------------------------------------------------------
// rows: one wave per root: first/last rows of the merged right vectors.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dc_rows(DcBufs B, int ibase) {
  DC_ITEM();
  const int K = B.ms[mi].kept;
  const int i = item;
  if (i >= K) return;
  const double* d = B.kd + M.off;
  const double* zh = B.zh + M.off;
  const double* F = B.kF + M.off;
  const double* L = B.kL + M.off;
  const int o = B.org[M.off + i];
  const double ti = B.tau[M.off + i];
  const double dO = d[o];
  const double w = dO + ti;
  double nn = 0.0, fs = 0.0, ls = 0.0;
  if (K > 1) {
    for (int j = lane; j < K; j += 64) {
      double v = zh[j] / (((d[j] - dO) - ti) * (d[j] + w));
      nn += v * v;
      fs += F[j] * v;
      ls += L[j] * v;
    }
    nn = wave_sum(nn);
    fs = wave_sum(fs);
    ls = wave_sum(ls);
  }
  if (lane != 0) return;
  if (nn > 0.0 && isfinite(nn)) {
    double inv = 1.0 / sqrt(nn);
    B.fn[M.off + i] = fs * inv;
    B.ln[M.off + i] = ls * inv;
  } else {
    // a single kept pole (the arrow is 1 x 1) or a vector lost to
    // underflow: the singular vector is the pole's own unit vector
    B.fn[M.off + i] = F[o];
    B.ln[M.off + i] = L[o];
  }
}

// ---------------------------------------------------------------------------
// assemble: merge the ascending roots with the ascending deflated values.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dc_assemble(DcBufs B, int mbase) {
  const int mi = mbase + blockIdx.x;
  const DcMerge M = B.merges[mi];
  const DcMergeState st = B.ms[mi];
  const int off = M.off, N = M.N, K = st.kept, ND = N - K;
  const double* d = B.kd + off;
  const double* dd = B.dd + off;
  // roots: w_i = (d_o + tau_i) * scale
  for (int i = threadIdx.x; i < K; i += 256) {
    double w = (d[B.org[off + i]] + B.tau[off + i]) * st.scale;
    int pos = i + count_less(dd, ND, w);
    if (M.top) {
      B.out[pos] = w;
    } else {
      B.sig[off + pos] = w;
      B.fr[off + pos] = B.fn[off + i];
      B.lr[off + pos] = B.ln[off + i];
    }
  }
  for (int q = threadIdx.x; q < ND; q += 256) {
    double v = dd[q];
    // ties go after the roots that equal v
    int lo = 0, hi = K;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      double wm = (d[B.org[off + mid]] + B.tau[off + mid]) * st.scale;
      if (wm <= v) lo = mid + 1; else hi = mid;
    }
    int pos = q + lo;
    if (M.top) {
      B.out[pos] = v;
    } else {
      B.sig[off + pos] = v;
      B.fr[off + pos] = B.dF[off + q];
      B.lr[off + pos] = B.dL[off + q];
    }
  }
}

// ---------------------------------------------------------------------------
// host planning
// ---------------------------------------------------------------------------
static int build_tree(std::vector<DcNode>& nodes, int off, int N) {
  int id = (int)nodes.size();
  nodes.push_back(DcNode{off, N, 0, -1, -1, 0});
  if (N >= 2) {
    int K = N / 2;
    int c1 = build_tree(nodes, off, K);
    int c2 = build_tree(nodes, off + K + 1, N - K - 1);
    nodes[id].K = K;
    nodes[id].c1 = c1;
    nodes[id].c2 = c2;
    nodes[id].height = 1 + std::max(nodes[c1].height, nodes[c2].height);
  }
  return id;
}

// Upper bound on the work items: per level at most n/4 + (merges of the
// level) groups, at most log2(n) + 2 levels.
static size_t dc_item_cap(size_t nn) {
  size_t lv = 2;
  while (((size_t)1 << lv) < nn) ++lv;
  return (nn / 4 + 2) * (lv + 2) + nn + 64;
}

void dc_plan(ArenaPlan& p, int n) {
  size_t nn = (size_t)n + 1;
  p.add<double>(nn * 3);          // sig fr lr
  p.add<double>(2 * (2 * nn + 2));  // phi psi
  p.add<double>(nn * 4);          // wd wz wF wL
  p.add<int>(nn);                 // flag
  p.add<double>(nn * 4);          // kd kz kF kL
  p.add<double>(nn * 3);          // dd dF dL
  p.add<int>(nn);                 // org
  p.add<double>(nn * 4);          // tau zh fn ln
  p.add<DcMergeState>(nn);
  p.add<DcMerge>(nn);
  p.add<int2>(dc_item_cap(nn));
  p.add<int3>(nn + 1);
}

int dc_run(int n, const double* a, const double* b, double* sigma, Arena& ws,
           hipStream_t st) {
  if (n < 1) return DDC_E_ARG;
  std::vector<DcNode> nodes;
  nodes.reserve(2 * (size_t)n + 2);
  build_tree(nodes, 0, n);
  const int H = nodes[0].height;

  // merges grouped by height, work items per level
  std::vector<DcMerge> merges;
  std::vector<int> mstart(H + 2, 0);
  std::vector<int3> leaves;
  for (int h = 1; h <= H; ++h) {
    mstart[h] = (int)merges.size();
    for (size_t id = 0; id < nodes.size(); ++id) {
      const DcNode& d = nodes[id];
      if (d.N >= 2 && d.height == h)
        merges.push_back(DcMerge{d.off, d.N, d.K, (int)id, d.c1, d.c2, id == 0 ? 1 : 0});
    }
  }
  mstart[H + 1] = (int)merges.size();
  for (size_t id = 0; id < nodes.size(); ++id)
    if (nodes[id].N < 2) leaves.push_back(int3{nodes[id].off, nodes[id].N, (int)id});
  std::vector<int2> items;
  std::vector<int> istart(H + 2, 0);
  for (int h = 1; h <= H; ++h) {
    istart[h] = (int)items.size();
    for (int mi = mstart[h]; mi < mstart[h + 1]; ++mi)
      for (int i0 = 0; i0 < merges[mi].N; i0 += 4) items.push_back(int2{mi, i0});
  }
  istart[H + 1] = (int)items.size();

  size_t nn = (size_t)n + 1;
  DcBufs B;
  B.a = const_cast<double*>(a);
  B.b = const_cast<double*>(b);
  B.sig = ws.take<double>(nn * 3);
  B.fr = B.sig + nn;
  B.lr = B.fr + nn;
  B.phi = ws.take<double>(2 * (2 * nn + 2));
  B.psi = B.phi + (2 * nn + 2);
  B.wd = ws.take<double>(nn * 4);
  B.wz = B.wd + nn;
  B.wF = B.wz + nn;
  B.wL = B.wF + nn;
  B.flag = ws.take<int>(nn);
  B.kd = ws.take<double>(nn * 4);
  B.kz = B.kd + nn;
  B.kF = B.kz + nn;
  B.kL = B.kF + nn;
  B.dd = ws.take<double>(nn * 3);
  B.dF = B.dd + nn;
  B.dL = B.dF + nn;
  B.org = ws.take<int>(nn);
  B.tau = ws.take<double>(nn * 4);
  B.zh = B.tau + nn;
  B.fn = B.zh + nn;
  B.ln = B.fn + nn;
  B.ms = ws.take<DcMergeState>(nn);
  DcMerge* dmerges = ws.take<DcMerge>(nn);
  int2* ditems = ws.take<int2>(dc_item_cap(nn));
  int3* dleaves = ws.take<int3>(nn + 1);
  if (!dleaves) return DDC_E_NOMEM;
  if (merges.size() > nn || items.size() > dc_item_cap(nn) || leaves.size() > nn + 1)
    return DDC_E_ARG;
  B.merges = dmerges;
  B.items = ditems;
  B.out = sigma;

  if (!merges.empty())
    DDC_HIP(hipMemcpyAsync(dmerges, merges.data(), sizeof(DcMerge) * merges.size(),
                           hipMemcpyHostToDevice, st));
  if (!items.empty())
    DDC_HIP(hipMemcpyAsync(ditems, items.data(), sizeof(int2) * items.size(),
                           hipMemcpyHostToDevice, st));
  DDC_HIP(hipMemcpyAsync(dleaves, leaves.data(), sizeof(int3) * leaves.size(),
                         hipMemcpyHostToDevice, st));
  // the host vectors must outlive the asynchronous copies
  DDC_HIP(hipStreamSynchronize(st));

  hipLaunchKernelGGL(dc_leaves, dim3((unsigned)ceil_div((int64_t)leaves.size(), 256)),
                     dim3(256), 0, st, B, dleaves, (int)leaves.size());
  DDC_LAUNCHED();
  if (H == 0) {  // n == 1
    DDC_HIP(hipMemcpyAsync(sigma, B.sig, sizeof(double), hipMemcpyDeviceToDevice, st));
    return DDC_OK;
  }
  for (int h = 1; h <= H; ++h) {
    const int nm = mstart[h + 1] - mstart[h];
    const int ni = istart[h + 1] - istart[h];
    hipLaunchKernelGGL(dc_prep, dim3(nm), dim3(256), 0, st, B, mstart[h]);
    DDC_LAUNCHED();
    hipLaunchKernelGGL(dc_roots, dim3(ni), dim3(256), 0, st, B, istart[h]);
    DDC_LAUNCHED();
    if (h < H) {
      hipLaunchKernelGGL(dc_zhat, dim3(ni), dim3(256), 0, st, B, istart[h]);
      DDC_LAUNCHED();
      hipLaunchKernelGGL(dc_rows, dim3(ni), dim3(256), 0, st, B, istart[h]);
      DDC_LAUNCHED();
    }
    hipLaunchKernelGGL(dc_assemble, dim3(nm), dim3(256), 0, st, B, mstart[h]);
    DDC_LAUNCHED();
  }
  return DDC_OK;
}

}  // namespace ddc
