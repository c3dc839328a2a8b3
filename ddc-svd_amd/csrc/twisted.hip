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
  const double* qd;   // [2][3][n]: rep 0 = B, rep 1 = reversed; d, ld, lld
  const double* sig_in;  // D&C values
  double* sig_out;       // refined values
  double* Vt;         // v_i[j] at j*ldv + i
  double* Ut;         // u_i[j] at j*ldv + i
  int ldv;
  double* Lp;         // scratch [n][T]
  double* S;          // scratch [n][T]
  double* Um;         // scratch [n][T]
  int T;              // scratch stride (2n)
  double smax;
};

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

// number of eigenvalues of L D L^T below lam
__device__ int tw_negcount(const double* d, const double* lld, int n, double lam,
                           double pivmin) {
  double s = -lam;
  int cnt = 0;
  for (int i = 0; i < n - 1; ++i) {
    double dp = guard(d[i] + s, pivmin);
    cnt += (dp < 0.0) ? 1 : 0;
    s = lld[i] * s / dp - lam;
  }
  double dp = guard(d[n - 1] + s, pivmin);
  cnt += (dp < 0.0) ? 1 : 0;
  return cnt;
}

__global__ __launch_bounds__(64) void tw_solve(TwArgs a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.n;
  if (tid >= 2 * n) return;
  const int rep = (tid >= n) ? 1 : 0;
  const int idx = tid - rep * n;
  const double* d = a.qd + (size_t)rep * 3 * n;
  const double* ld = d + n;
  const double* lld = d + 2 * (size_t)n;
  const int64_t T = a.T;
  double* Lp = a.Lp + tid;
  double* S = a.S + tid;
  double* Um = a.Um + tid;

  const double smax = a.smax;
  const double pivmin = TW_TINY * fmax(1.0, smax * smax);
  const double sig = a.sig_in[idx];
  double lam = sig * sig;
  const double dl = 8.0 * n * TW_EPS * smax;
  double lo = fmax(0.0, sig - dl);
  lo *= lo;
  double hi = (sig + dl) * (sig + dl);
  if (tw_negcount(d, lld, n, lo, pivmin) > idx) lo = 0.0;
  if (tw_negcount(d, lld, n, hi, pivmin) < idx + 1) hi = (smax + dl) * (smax + dl) * (1.0 + 1e-8);

  double* out = rep ? a.Ut : a.Vt;
  double nz2 = 1.0;
  for (int it = 0; it < 24; ++it) {
    // stationary transform: L D L^T - lam = L+ D+ L+^T
    double s = -lam;
    for (int i = 0; i < n - 1; ++i) {
      S[i * T] = s;
      double dp = guard(d[i] + s, pivmin);
      Lp[i * T] = ld[i] / dp;
      s = lld[i] * s / dp - lam;
    }
    S[(n - 1) * T] = s;
    // progressive transform: = U- D- U-^T, twist values gamma_i = s_i + p_i + lam
    double p = d[n - 1] - lam;
    double gbest = s + p + lam;
    int r = n - 1;
    for (int i = n - 2; i >= 0; --i) {
      double dm = guard(lld[i] + p, pivmin);
      Um[i * T] = ld[i] / dm;
      p = d[i] * p / dm - lam;
      double g = S[i * T] + p + lam;
      if (fabs(g) < fabs(gbest)) {
        gbest = g;
        r = i;
      }
    }
    // z with z_r = 1
    auto put = [&](int j, double v) {
      int jj = rep ? (n - 1 - j) : j;
      out[(int64_t)jj * a.ldv + idx] = v;
    };
    double z = 1.0;
    nz2 = 1.0;
    put(r, 1.0);
    for (int i = r; i < n - 1; ++i) {
      z = -Um[i * T] * z;
      nz2 += z * z;
      put(i + 1, z);
    }
    z = 1.0;
    for (int i = r - 1; i >= 0; --i) {
      z = -Lp[i * T] * z;
      nz2 += z * z;
      put(i, z);
    }
    double corr = gbest / nz2;
    if (fabs(corr) <= 4.0 * TW_EPS * lam || nz2 > 1.0 / (TW_EPS * TW_EPS)) break;
    double nl = lam + corr;
    if (!(nl > lo && nl < hi)) nl = 0.5 * (lo + hi);
    if (tw_negcount(d, lld, n, nl, pivmin) <= idx) lo = nl; else hi = nl;
    lam = nl;
  }
  const double inv = 1.0 / sqrt(nz2);
  for (int j = 0; j < n; ++j) out[(int64_t)j * a.ldv + idx] *= inv;
  if (rep == 0) a.sig_out[idx] = sqrt(fmax(lam, 0.0));
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
  hipLaunchKernelGGL(tw_solve, dim3((unsigned)ceil_div(2 * n, 64)), dim3(64), 0, st, t);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(tw_sign, dim3((unsigned)ceil_div(n, 64)), dim3(64), 0, st, n, a, b, Vt, Ut, ldv);
  DDC_LAUNCHED();
  return DDC_OK;
}

}  // namespace ddc
