// twisted.hip -- per-singular-value vectors by twisted factorization, FP64.
//
// Replaces CalcRightSingularVectors and RighttoLeftSingularVectors
// (parallel-twisted.c:530-637).  The reference factors B^T B - sigma^2 I
// after forming B^T B (SquareB :290-316), which loses the small singular
// values, and builds u = B v / sigma, which divides by zero for sigma = 0.
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
// neighbouring singular value.  The divide-and-conquer value seeds the
// iteration; on exit sigma holds the refined value (high relative accuracy).
//
// Left vectors are computed the same way from B B^T, which is the Gram
// matrix of the reversed bidiagonal B' = J B^T J (diagonal a reversed,
// super-diagonal b reversed); u = J z'.  The sign of u is then fixed so that
// u^T B v >= 0.
#include "ddc_common.h"

namespace ddc {

constexpr double TW_EPS = 2.220446049250313e-16;
constexpr double TW_TINY = 2.2250738585072014e-308;

struct TwArgs {
  int n;
  const double* qd;      // [2][3][n]: rep 0 = B, rep 1 = reversed; d, ld, lld
  const double* sig_in;  // D&C values
  double* sig_out;       // refined values
  double* Vt;            // v_i[j] at j*ldv + i
  double* Ut;            // u_i[j] at j*ldv + i
  int ldv;
  double* Lp;            // scratch [n][T]: stationary multipliers L+
  double* S;             // scratch [n][T]: stationary s_i
  double* Um;            // scratch [n][T]: progressive multipliers U-
  int T;                 // scratch stride (2n)
  double smax;
};

constexpr int TW_CH = 256;  // qd entries staged in LDS per chunk

// qd data for B and its reversal
__global__ void tw_qd(int n, const double* a, const double* b, double* qd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* d0 = qd;
  double* ld0 = qd + n;
  double* lld0 = qd + 2 * (size_t)n;
  double* d1 = qd + 3 * (size_t)n;
  double* ld1 = qd + 4 * (size_t)n;
  double* lld1 = qd + 5 * (size_t)n;
  d0[i] = a[i] * a[i];
  int r = n - 1 - i;
  d1[i] = a[r] * a[r];
  if (i < n - 1) {
    ld0[i] = a[i] * b[i];
    lld0[i] = b[i] * b[i];
    // reversed: a'_i = a[n-1-i], b'_i = b[n-2-i]
    ld1[i] = a[r] * b[r - 1];
    lld1[i] = b[r - 1] * b[r - 1];
  } else {
    ld0[i] = lld0[i] = ld1[i] = lld1[i] = 0.0;
  }
}

__device__ __forceinline__ double guard(double x, double pivmin) {
  return (fabs(x) < pivmin) ? -pivmin : x;
}

// One wave per group, all lanes on the same representation; lane = one
// singular value.  All lanes walk the index range in lockstep, so the qd
// arrays (shared by the lanes) are staged through LDS chunk by chunk and the
// per-lane scratch accesses (column-interleaved, stride T) are coalesced.
__global__ __launch_bounds__(64) void tw_solve(TwArgs a) {
  __shared__ double fd[TW_CH], fld[TW_CH], flld[TW_CH];  // ascending chunk
  __shared__ double bd[TW_CH], bld[TW_CH], blld[TW_CH];  // descending chunk
  const int n = a.n;
  const int nbr = (n + 63) / 64;
  const int rep = ((int)blockIdx.x >= nbr) ? 1 : 0;
  const int lane = threadIdx.x;
  const int idx = ((int)blockIdx.x - rep * nbr) * 64 + lane;
  const bool live = idx < n;
  const double* __restrict__ d = a.qd + (size_t)rep * 3 * n;
  const double* __restrict__ ld = d + n;
  const double* __restrict__ lld = d + 2 * (size_t)n;
  const int64_t T = a.T;
  const int col = rep * n + (live ? idx : 0);
  double* __restrict__ Lp = a.Lp + col;
  double* __restrict__ S = a.S + col;
  double* __restrict__ Um = a.Um + col;
  double* __restrict__ out = (rep ? a.Ut : a.Vt) + (live ? idx : 0);
  const int64_t ldv = a.ldv;
  auto oidx = [&](int j) -> int64_t { return (int64_t)(rep ? (n - 1 - j) : j) * ldv; };

  const double smax = a.smax;
  const double pivmin = TW_TINY * fmax(1.0, smax * smax);
  const double sig = live ? a.sig_in[idx] : 0.0;
  const double dl = 8.0 * n * TW_EPS * smax;
  double lo = fmax(0.0, sig - dl);
  lo *= lo;
  double hi = (sig + dl) * (sig + dl);
  const double dn1 = d[n - 1];

  // Sturm counts at lo and hi in one ascending pass (two independent chains)
  {
    double s1 = -lo, s2 = -hi;
    int c1 = 0, c2 = 0;
    for (int c0 = 0; c0 < n - 1; c0 += TW_CH) {
      const int len = min(TW_CH, n - 1 - c0);
      __syncthreads();
      for (int e = lane; e < len; e += 64) {
        fd[e] = d[c0 + e];
        flld[e] = lld[c0 + e];
      }
      __syncthreads();
      for (int e = 0; e < len; ++e) {
        double p1 = guard(fd[e] + s1, pivmin);
        double p2 = guard(fd[e] + s2, pivmin);
        c1 += (p1 < 0.0);
        c2 += (p2 < 0.0);
        s1 = flld[e] * s1 / p1 - lo;
        s2 = flld[e] * s2 / p2 - hi;
      }
    }
    c1 += (guard(dn1 + s1, pivmin) < 0.0);
    c2 += (guard(dn1 + s2, pivmin) < 0.0);
    if (c1 > idx) lo = 0.0;
    if (c2 < idx + 1) hi = (smax + dl) * (smax + dl) * (1.0 + 1e-8);
  }

  double lam = sig * sig;
  double nz2 = 1.0, corr = 0.0;
  bool active = live;
  for (int it = 0; it < 24; ++it) {
    if (__ballot(active) == 0) break;
    // stationary (ascending, stores s_i and L+_i, counts negative pivots)
    // and progressive (descending, stores p_i and U-_i) transforms in lockstep
    double s = -lam;
    double p = dn1 - lam;
    int cnt = 0;
    if (active) out[(int64_t)(n - 1) * ldv] = p;  // p_{n-1}
    for (int c0 = 0; c0 < n - 1; c0 += TW_CH) {
      const int len = min(TW_CH, n - 1 - c0);
      const int g0 = n - 1 - c0 - len;  // descending chunk covers [g0, g0+len)
      __syncthreads();
      for (int e = lane; e < len; e += 64) {
        fd[e] = d[c0 + e];
        fld[e] = ld[c0 + e];
        flld[e] = lld[c0 + e];
        bd[e] = d[g0 + e];
        bld[e] = ld[g0 + e];
        blld[e] = lld[g0 + e];
      }
      __syncthreads();
      if (active) {
        for (int e = 0; e < len; ++e) {
          const int f = c0 + e;
          const int gi = len - 1 - e;  // descending: g = g0 + gi
          const int64_t g = g0 + gi;
          S[(int64_t)f * T] = s;
          double dp = guard(fd[e] + s, pivmin);
          cnt += (dp < 0.0);
          Lp[(int64_t)f * T] = fld[e] / dp;
          s = flld[e] * s / dp - lam;
          double dm = guard(blld[gi] + p, pivmin);
          Um[g * T] = bld[gi] / dm;
          p = bd[gi] * p / dm - lam;
          out[g * ldv] = p;  // p_g (scratch use of the output column)
        }
      }
    }
    if (!active) continue;
    S[(int64_t)(n - 1) * T] = s;
    cnt += (guard(dn1 + s, pivmin) < 0.0);
    // twist index: smallest |gamma_i| = |s_i + p_i + lam| (scan from the
    // top; ties keep the larger index)
    int r = n - 1;
    double gbest = s + out[(int64_t)(n - 1) * ldv] + lam;
    for (int i0 = n - 2; i0 >= 0; i0 -= 8) {
      double sv[8], pv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = max(i0 - u, 0);
        sv[u] = S[(int64_t)i * T];
        pv[u] = out[(int64_t)i * ldv];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double gm = sv[u] + pv[u] + lam;
        if (i0 - u >= 0 && fabs(gm) < fabs(gbest)) {
          gbest = gm;
          r = i0 - u;
        }
      }
    }
    // z with z_r = 1: downward z_{i+1} = -U-_i z_i (i >= r), upward
    // z_i = -L+_i z_{i+1} (i < r); both sweeps run over the whole range so
    // that the lanes stay at the same index
    nz2 = 1.0;
    double z = 1.0;
    out[oidx(r)] = 1.0;
    for (int i0 = 0; i0 < n - 1; i0 += 8) {
      double uv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) uv[u] = Um[(int64_t)min(i0 + u, n - 2) * T];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u;
        if (i < n - 1 && i >= r) {
          z = -uv[u] * z;
          nz2 += z * z;
          out[oidx(i + 1)] = z;
        }
      }
    }
    z = 1.0;
    for (int i0 = n - 2; i0 >= 0; i0 -= 8) {
      double lv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) lv[u] = Lp[(int64_t)max(i0 - u, 0) * T];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 - u;
        if (i >= 0 && i < r) {
          z = -lv[u] * z;
          nz2 += z * z;
          out[oidx(i)] = z;
        }
      }
    }
    corr = gbest / nz2;
    if (fabs(corr) <= 4.0 * TW_EPS * lam || nz2 > 1.0 / (TW_EPS * TW_EPS)) {
      active = false;
      continue;
    }
    // cnt = number of eigenvalues below lam: keep the bracket around idx
    if (cnt <= idx) lo = lam; else hi = lam;
    double nl = lam + corr;
    if (!(nl > lo && nl < hi)) nl = 0.5 * (lo + hi);
    lam = nl;
  }
  if (!live) return;
  const double inv = 1.0 / sqrt(nz2);
  for (int j0 = 0; j0 < n; j0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = out[(int64_t)min(j0 + u, n - 1) * ldv];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < n) out[(int64_t)(j0 + u) * ldv] = v[u] * inv;
  }
  if (rep == 0) a.sig_out[idx] = sqrt(fmax(lam + corr, 0.0));
}

// u_i^T B v_i >= 0
__global__ void tw_sign(int n, const double* a, const double* b, const double* Vt,
                        double* Ut, int ldv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = 0.0;
  for (int j = 0; j < n; ++j) {
    double bv = a[j] * Vt[(int64_t)j * ldv + i];
    if (j < n - 1) bv += b[j] * Vt[(int64_t)(j + 1) * ldv + i];
    t += Ut[(int64_t)j * ldv + i] * bv;
  }
  if (t < 0.0)
    for (int j = 0; j < n; ++j) Ut[(int64_t)j * ldv + i] = -Ut[(int64_t)j * ldv + i];
}

void twisted_plan(ArenaPlan& p, int n) {
  p.add<double>(6 * (size_t)n);
  p.add<double>(3 * (size_t)n * 2 * n);
  p.add<double>(1);
  p.add<double>(n);
}

int twisted_run(int n, const double* a, const double* b, double* sigma,
                double* Vt, double* Ut, int ldv, Arena& ws, hipStream_t st) {
  if (n < 1 || ldv < n) return DDC_E_ARG;
  double* qd = ws.take<double>(6 * (size_t)n);
  double* scr = ws.take<double>(3 * (size_t)n * 2 * n);
  double* hsmax = ws.take<double>(1);
  (void)hsmax;
  double* sig0 = ws.take<double>(n);
  if (!sig0) return DDC_E_NOMEM;
  DDC_HIP(hipMemcpyAsync(sig0, sigma, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(tw_qd, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n, a, b, qd);
  DDC_LAUNCHED();
  double smax = 0.0;
  DDC_HIP(hipMemcpyAsync(&smax, sigma + (n - 1), sizeof(double), hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  TwArgs t;
  t.n = n;
  t.qd = qd;
  t.sig_in = sig0;
  t.sig_out = sigma;
  t.Vt = Vt;
  t.Ut = Ut;
  t.ldv = ldv;
  t.T = 2 * n;
  t.Lp = scr;
  t.S = scr + (size_t)n * 2 * n;
  t.Um = scr + 2 * (size_t)n * 2 * n;
  t.smax = smax;
  hipLaunchKernelGGL(tw_solve, dim3(2 * (unsigned)ceil_div(n, 64)), dim3(64), 0, st, t);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(tw_sign, dim3((unsigned)ceil_div(n, 64)), dim3(64), 0, st, n, a, b, Vt, Ut, ldv);
  DDC_LAUNCHED();
  return DDC_OK;
}

}  // namespace ddc
