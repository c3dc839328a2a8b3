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
  double* Sck;           // checkpoints of s every TW_CK steps [nck][T]
  double* Pck;           // checkpoints of p every TW_CK steps [nck][T]
  int T;                 // checkpoint stride (2n)
  double smax;
};

constexpr int TW_CK = 64;  // checkpoint interval = LDS chunk

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
//            and its argmin over the chunk's two halves
//   pass 3   wave 0: z downward from the twist (recomputed p); wave 1: z
//            upward (recomputed s)
//
// Recurrences (L D L^T - lam I, d = a^2, ld = a b, lld = b^2):
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
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int idx = ((int)blockIdx.x - rep * nbr) * 64 + lane;
  const bool live = idx < n;
  const double* d = a.qd + (size_t)rep * 3 * n;
  const double* ld = d + n;
  const double* lld = d + 2 * (size_t)n;
  const int64_t T = a.T;
  const int col = rep * n + (live ? idx : 0);
  double* Sck = a.Sck + col;
  double* Pck = a.Pck + col;
  double* out = (rep ? a.Ut : a.Vt) + (live ? idx : 0);
  const int64_t ldv = a.ldv;
  auto oidx = [&](int j) -> int64_t { return (int64_t)(rep ? (n - 1 - j) : j) * ldv; };
  const int nck = (n + TW_CK - 1) / TW_CK;
  // both waves stage one chunk each (cw: this wave's chunk)
  auto stage = [&](int cw) {
    __syncthreads();
    const int c0 = cw * TW_CK, len = min(TW_CK, n - c0);
    if (lane < len) {
      qs[wave][0][lane] = d[c0 + lane];
      qs[wave][1][lane] = ld[c0 + lane];
      qs[wave][2][lane] = lld[c0 + lane];
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
  const double dl = 8.0 * n * TW_EPS * smax;
  double lo = fmax(0.0, sig - dl);
  lo *= lo;
  double hi = (sig + dl) * (sig + dl);
  const double dn1 = d[n - 1];

  // Sturm counts: wave 0 at lo, wave 1 at hi
  {
    const double sh = wave == 0 ? lo : hi;
    double s = -sh;
    int cn = 0;
    for (int c = 0; c < nck; ++c) {
      stage(c);
      const int len = min(TW_CK, n - c * TW_CK);
#pragma unroll 8
      for (int e = 0; e < TW_CK; ++e) {
        if (e < len) {
          const double p = guard(qd_d[e] + s, pivmin);
          cn += (p < 0.0);
          s = qd_lld[e] * s / p - sh;
        }
      }
    }
    xr[wave][lane] = cn;
    __syncthreads();
    const int c1 = xr[0][lane], c2 = xr[1][lane];
    if (c1 > idx) lo = 0.0;
    if (c2 < idx + 1) hi = (smax + dl) * (smax + dl) * (1.0 + 1e-8);
  }

  double lam = sig * sig;
  double nz2 = 1.0, corr = 0.0, prev = INFINITY;
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
#pragma unroll 8
          for (int e = 0; e < TW_CK; ++e) {
            if (e < len) {
              const double dp = guard(qd_d[e] + s, pivmin);
              cnt += (dp < 0.0);
              s = qd_lld[e] * s / dp - lam;  // past the last index: unused
            }
          }
        } else {
#pragma unroll 8
          for (int e = TW_CK - 1; e >= 0; --e) {
            if (e < len && c0 + e < n - 1) {
              const double dm = guard(qd_lld[e] + p, pivmin);
              p = qd_d[e] * p / dm - lam;
            }
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
#pragma unroll 8
        for (int e = 0; e < TW_CK; ++e) {
          if (e < len) {
            bw[e][lane] = ss;
            const double dp = guard(qd_d[e] + ss, pivmin);
            ss = qd_lld[e] * ss / dp - lam;
          }
        }
      } else {
        double pp = (c + 1 < nck) ? Pck[(int64_t)(c + 1) * T] : 0.0;  // p_{c0+len}
#pragma unroll 8
        for (int e = TW_CK - 1; e >= 0; --e) {
          if (e < len) {
            if (c0 + e == n - 1) {
              pp = dn1 - lam;
            } else {
              const double dm = guard(qd_lld[e] + pp, pivmin);
              pp = qd_d[e] * pp / dm - lam;
            }
            bw[e][lane] = pp;
          }
        }
      }
      __syncthreads();
      // wave 0 takes e < 32 of the chunk, wave 1 e >= 32
#pragma unroll 4
      for (int q = 0; q < TW_CK / 2; ++q) {
        const int e = wave * (TW_CK / 2) + q;
        if (e < len) {
          const double g = buf[0][e][lane] + buf[1][e][lane] + lam;
          if (fabs(g) < fabs(gbest) || (fabs(g) == fabs(gbest) && fabs(g) < INFINITY)) {
            gbest = g;
            r = c0 + e;
          }
        }
      }
    }
    // combine: the smaller |gamma|, the larger index on ties; with no
    // finite gamma at all the twist stays at the last index
    {
      xg[wave][lane] = gbest;
      xr[wave][lane] = r;
      if (wave == 0) xc[lane] = cnt;
      __syncthreads();
      const double g0 = xg[0][lane], g1 = xg[1][lane];
      const int r0 = xr[0][lane], r1 = xr[1][lane];
      cnt = xc[lane];
      const bool take1 =
          r1 >= 0 && (r0 < 0 || fabs(g1) < fabs(g0) || (fabs(g1) == fabs(g0) && r1 > r0));
      if (take1) {
        gbest = g1;
        r = r1;
      } else if (r0 >= 0) {
        gbest = g0;
        r = r0;
      } else {
        gbest = INFINITY;
        r = n - 1;
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
#pragma unroll 8
          for (int e = TW_CK - 1; e >= 0; --e) {
            if (e <= etop) {
              bw[e][lane] = pp;  // p_{c0+e+1}
              const double dm = guard(qd_lld[e] + pp, pivmin);
              pp = qd_d[e] * pp / dm - lam;
            }
          }
#pragma unroll 8
          for (int e = 0; e < TW_CK; ++e) {
            const int i = c0 + e;
            if (e <= etop && active && i >= r) {
              z = -(qd_ld[e] / guard(qd_lld[e] + bw[e][lane], pivmin)) * z;
              nzp += z * z;
              out[oidx(i + 1)] = z;
            }
          }
        } else {
          double ss = Sck[(int64_t)cw * T];
#pragma unroll 8
          for (int e = 0; e < TW_CK; ++e) {
            if (e < len) {
              bw[e][lane] = ss;
              const double dp = guard(qd_d[e] + ss, pivmin);
              ss = qd_lld[e] * ss / dp - lam;
            }
          }
#pragma unroll 8
          for (int e = TW_CK - 1; e >= 0; --e) {
            const int i = c0 + e;
            if (e < len && active && i < r) {
              z = -(qd_ld[e] / guard(qd_d[e] + bw[e][lane], pivmin)) * z;
              nzp += z * z;
              out[oidx(i)] = z;
            }
          }
        }
      }
    }
    xn[wave][lane] = nzp;
    __syncthreads();
    // converged lanes keep the norm of their final vector
    if (active) nz2 = xn[0][lane] + xn[1][lane];
    if (!active) continue;
    corr = gbest / nz2;
    // converged: correction at the rounding level of lam, the vector's norm
    // saturated, or (from the third step on) the correction no longer
    // halving -- it has reached the rounding floor of gamma_r, which for
    // small singular values sits above eps*lam
    if (fabs(corr) <= 4.0 * TW_EPS * lam || nz2 > 1.0 / (TW_EPS * TW_EPS) ||
        (it >= 2 && fabs(corr) > 0.5 * prev)) {
      active = false;
      continue;
    }
    prev = fabs(corr);
    // cnt = number of eigenvalues below lam: keep the bracket around idx
    if (cnt <= idx) lo = lam; else hi = lam;
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
  if (rep == 0 && wave == 0) a.sig_out[idx] = sqrt(fmax(lam + corr, 0.0));
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

static size_t tw_ck_size(int n) { return (size_t)((n + TW_CK - 1) / TW_CK) * 2 * n; }

void twisted_plan(ArenaPlan& p, int n) {
  p.add<double>(6 * (size_t)n);
  p.add<double>(2 * tw_ck_size(n));
  p.add<double>(1);
  p.add<double>(n);
}

int twisted_run(int n, const double* a, const double* b, double* sigma,
                double* Vt, double* Ut, int ldv, Arena& ws, hipStream_t st) {
  if (n < 1 || ldv < n) return DDC_E_ARG;
  double* qd = ws.take<double>(6 * (size_t)n);
  double* scr = ws.take<double>(2 * tw_ck_size(n));
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
  t.Sck = scr;
  t.Pck = scr + tw_ck_size(n);
  t.smax = smax;
  hipLaunchKernelGGL(tw_solve, dim3(2 * (unsigned)ceil_div(n, 64)), dim3(128), 0, st, t);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(tw_sign, dim3((unsigned)ceil_div(n, 64)), dim3(256), 0, st, n, a, b, Vt,
                     Ut, ldv);
  DDC_LAUNCHED();
  return DDC_OK;
}

}  // namespace ddc
