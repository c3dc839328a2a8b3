// adjoint.hip -- reverse-mode derivative of the thin SVD, FP64 on the GPU.
//
// No reference counterpart (the reference computes the forward SVD only).
// For A = U S V^T (m >= n, U m x n, V n x n, distinct singular values) and a
// scalar loss L with cotangents Ubar = dL/dU, sbar = dL/dsigma, Vbar = dL/dV:
//
//   Abar = U [ (J o F) S + S (K o F) + diag(sbar) ] V^T + (I - U U^T) Ubar S^-1 V^T
//   J = U^T Ubar - Ubar^T U,  K = V^T Vbar - Vbar^T V,
//   F_ij = 1 / (s_j^2 - s_i^2) (i != j),  F_ii = 0.
//
// Rewritten for the GPU as  Abar = (U W + Ubar S^-1) V^T  with
//   W = (J o F) S + S (K o F) + diag(sbar) - (U^T Ubar) S^-1,
// which costs four library GEMMs of the thin size (U^T Ubar, V^T Vbar and the
// two products, FP64 MFMA) plus one elementwise "F-matrix" kernel.  The gap guard
// replaces 1/d by d/delta^2 when |d| = |s_j^2 - s_i^2| < delta =
// guard * s_max^2 (continuous at |d| = delta; guard 0 gives the plain
// formula), and 1/s_j by 0 for s_j <= guard * s_max.
#include <algorithm>

#include "ddc_common.h"

namespace ddc {

struct AdjArgs {
  int n;
  const double* sigma;
  const double* P;     // U^T Ubar (n x n, ld n), or null
  const double* Q;     // V^T Vbar (n x n, ld n), or null
  const double* sbar;  // or null
  double* W;           // n x n, ld n
  double delta2;       // (guard * s_max^2)^2
  double smin;         // singular values at or below this count as zero
};

__device__ __forceinline__ double guarded_inv(double d, double delta2) {
  const double d2 = d * d;
  if (d2 >= delta2 && d != 0.0) return 1.0 / d;
  return (delta2 > 0.0) ? d / delta2 : 0.0;
}

// One 32 x 32 tile of W per group; the transposed tiles of P and Q come
// through LDS so that both reads are coalesced.
__global__ __launch_bounds__(256) void adj_fmatrix(AdjArgs a) {
  __shared__ double Pt[32][33], Qt[32][33];
  const int n = a.n;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  // transposed tiles: Pt[c][r] = P(j0 + r, i0 + c)
  for (int r = ty; r < 32; r += 8) {
    const int jj = j0 + r, ii = i0 + tx;
    const bool ok = jj < n && ii < n;
    Pt[tx][r] = (ok && a.P) ? a.P[jj + (int64_t)ii * n] : 0.0;
    Qt[tx][r] = (ok && a.Q) ? a.Q[jj + (int64_t)ii * n] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 32; c += 8) {
    const int i = i0 + tx, j = j0 + c;
    if (i >= n || j >= n) continue;
    const double si = a.sigma[i], sj = a.sigma[j];
    const double pij = a.P ? a.P[i + (int64_t)j * n] : 0.0;
    const double qij = a.Q ? a.Q[i + (int64_t)j * n] : 0.0;
    const double pji = Pt[tx][c], qji = Qt[tx][c];
    double w = 0.0;
    if (i != j) {
      const double f = guarded_inv((sj - si) * (sj + si), a.delta2);
      w = (pij - pji) * f * sj + si * (qij - qji) * f;
    } else if (a.sbar) {
      w = a.sbar[i];
    }
    if (sj > a.smin) w -= pij / sj;
    a.W[i + (int64_t)j * n] = w;
  }
}

// G = Ubar S^-1 (columns scaled), zero where Ubar is absent.
__global__ void adj_scale(int m, int n, const double* Ub, int ldub, const double* sigma,
                          double smin, double* G, int ldg) {
  const int64_t total = (int64_t)m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / m, i = e - j * m;
    const double s = sigma[j];
    G[i + j * ldg] = (Ub && s > smin) ? Ub[i + j * ldub] / s : 0.0;
  }
}

void adjoint_plan(ArenaPlan& p, int m, int n) {
  p.add<double>((size_t)n * n * 3);  // P, Q, W
  p.add<double>((size_t)m * n);      // G
}

// Tall case (m >= n).  All pointers on the device; null cotangents are zero.
int adjoint_run(int m, int n, const double* U, int ldu, const double* sigma,
                const double* V, int ldv, const double* Ub, int ldub, const double* sb,
                const double* Vb, int ldvb, double* Ab, int ldab, double guard,
                Arena& ws, hipStream_t st) {
  if (m < n || n < 1) return DDC_E_ARG;
  double* P = ws.take<double>((size_t)n * n * 3);
  double* Q = P + (size_t)n * n;
  double* W = Q + (size_t)n * n;
  double* G = ws.take<double>((size_t)m * n);
  if (!G) return DDC_E_NOMEM;
  double smax = 0.0;
  DDC_HIP(hipMemcpyAsync(&smax, sigma + (n - 1), sizeof(double), hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));

  if (Ub) DDC_TRY(dgemm('T', 'N', n, n, m, 1.0, U, ldu, Ub, ldub, 0.0, P, n, st));
  if (Vb) DDC_TRY(dgemm('T', 'N', n, n, n, 1.0, V, ldv, Vb, ldvb, 0.0, Q, n, st));
  AdjArgs a;
  a.n = n;
  a.sigma = sigma;
  a.P = Ub ? P : nullptr;
  a.Q = Vb ? Q : nullptr;
  a.sbar = sb;
  a.W = W;
  const double delta = guard * smax * smax;
  a.delta2 = delta * delta;
  a.smin = guard * smax;
  dim3 g((unsigned)ceil_div(n, 32), (unsigned)ceil_div(n, 32));
  hipLaunchKernelGGL(adj_fmatrix, g, dim3(256), 0, st, a);
  DDC_LAUNCHED();
  hipLaunchKernelGGL(adj_scale, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)m * n, 256), 4096)),
                     dim3(256), 0, st, m, n, Ub, ldub, sigma, a.smin, G, m);
  DDC_LAUNCHED();
  // G += U W ; Abar = G V^T
  DDC_TRY(dgemm('N', 'N', m, n, n, 1.0, U, ldu, W, n, 1.0, G, m, st));
  DDC_TRY(dgemm('N', 'T', m, n, n, 1.0, G, m, V, ldv, 0.0, Ab, ldab, st));
  return DDC_OK;
}

}  // namespace ddc
