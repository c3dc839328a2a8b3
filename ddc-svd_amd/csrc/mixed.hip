// mixed.hip -- the mixed-precision SVD (svd_gpu_mixed), square matrices.
//
// No reference counterpart (the reference computes in FP64 only,
// svd_gpu.c:53).  Stages:
//   1. A (FP64) -> FP32 copy; bidiagonalization with the matrix and its
//      reflectors stored in FP32 (halving the GEMV traffic, which bounds the
//      plain path), vector arithmetic in FP64, and the blocked trailing
//      update in FP32 or BF16 on the matrix cores with FP32 accumulation;
//   2. singular values and bidiagonal vectors in FP64 (divide and conquer,
//      twisted factorizations), back-transformation in FP32 on the matrix
//      cores (its rounding is of the order of the FP32 reflectors' own, and
//      the refinement removes both), U and V widened to FP64;
//   3. refinement steps in FP64 (Ogita and Aishima, "Iterative refinement
//      for singular value decomposition based on matrix multiplication",
//      J. Comput. Appl. Math. 369 (2020)): with T = U^T A V, R = I - U^T U,
//      S = I - V^T V,
//         sigma_i = T_ii / (1 - (R_ii + S_ii) / 2)
//      and for i != j the 2 x 2 system of the first-order conditions
//         F + F^T = R,  G + G^T = S,  T + F^T Sigma + Sigma G = Sigma
//      gives F_ij, G_ij; then U <- U (I + F), V <- V (I + G).
//      Each step is two FP64 GEMMs and two SYRKs (library, MFMA) for T, R
//      and S, and two FP32 GEMMs for the update products U F and V G;
//   4. optionally (resolve) an FP64 Rayleigh-Ritz step on the values whose
//      pairs the last step could not separate: where a spectrum's absolute
//      gaps fall below the FP32 stage's error (graded spectra) no number of
//      first-order steps separates those vectors, the FP64 SVD of their
//      projected block does.
#include <algorithm>
#include <vector>

#include "ddc_common.h"

namespace ddc {

// Largest first-order correction applied.  The default suits spectra with
// long runs of values closer than the FP32 stage's error (their vectors are
// only re-orthogonalised); the larger bound, chosen by the caller for
// spectra whose close values come in isolated pairs, corrects those pairs
// too, and a second step then brings them to FP64 accuracy.
constexpr double OA_MAXCORR = 1e-3;
constexpr double OA_MAXCORR_SEPARATED = 1e-2;

// B (rows x cols, ld ldb, float) = A (double); rows rows..ldb-1 are zeroed.
__global__ void mx_to_f32(int rows, int cols, const double* A, int lda, float* B, int ldb) {
  const int64_t total = (int64_t)ldb * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / ldb, i = e - j * ldb;
    B[e] = i < rows ? (float)A[i + j * lda] : 0.0f;
  }
}

// B (n x n, ld ldb, double) = A (n x n, ld n, float).
__global__ void mx_to_f64(int n, const float* A, double* B, int ldb) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n, i = e - j * n;
    B[i + j * ldb] = (double)A[e];
  }
}

// D (n x n, ld n, float) = S (n x n, ld n, double).
__global__ void mx_round_f32(int n, const double* S, float* D) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    D[e] = (float)S[e];
}

// Un (ld ldun) = Uc + UF for the n x n matrices Uc (double, ld n) and UF
// (float, ld n).
__global__ void mx_add_f32(int n, const double* Uc, const float* UF, double* Un, int ldun) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n, i = e - j * n;
    Un[i + j * ldun] = Uc[e] + (double)UF[e];
  }
}

// Partial sums of (delta_ij - M_ij)^2 over the n x n matrix M (ld n), one
// per work-group, in a fixed order.
__global__ __launch_bounds__(256) void mx_defect(int n, const double* M, double* part) {
  __shared__ double red[4];
  double s = 0.0;
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t j = e / n, i = e - j * n;
    const double d = (i == j ? 1.0 : 0.0) - M[e];
    s += d * d;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Lower triangle of the n x n matrix C (ld n) from its upper triangle, one
// 32 x 32 tile pair per work-group through LDS (both sides coalesced).
__global__ void mx_symm(int n, double* C) {
  __shared__ double tile[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;  // lower tile (bi, bj), bi >= bj
  if (bi < bj) return;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  // read the upper tile (bj, bi): rows bj*32.., columns bi*32..
  for (int q = ty; q < 32; q += 8) {
    const int r = bj * 32 + tx, c = bi * 32 + q;
    tile[q][tx] = (r < n && c < n) ? C[r + (int64_t)c * n] : 0.0;
  }
  __syncthreads();
  // write the lower tile (bi, bj): C(i, j) = C(j, i) for i > j
  for (int q = ty; q < 32; q += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + q;
    if (i < n && j < n && i > j) C[i + (int64_t)j * n] = tile[tx][q];
  }
}

// sigma_i = T_ii / (1 - (r_ii + s_ii)/2) with r = I - U^T U, s = I - V^T V
// (P = U^T U, Q = V^T V given).
__global__ void oa_sigma(int n, const double* T, const double* P, const double* Q, double* sig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = (int64_t)i * n + i;
  sig[i] = T[d] / (0.5 * (P[d] + Q[d]));
}

// F and G of the refinement (n x n, ld n).  Off the diagonal:
//   a = -t_ij - r_ij s_j,  b = -t_ji - s_j q_ij,  d = s_j^2 - s_i^2
//   f_ij = (-s_j a - s_i b) / d,  g_ij = (-s_i a - s_j b) / d
// with r = I - U^T U, q = I - V^T V; on it f_ii = r_ii / 2, g_ii = q_ii / 2.
// Pairs closer than the guard, and pairs whose correction is not small
// (|f_ij| or |g_ij| > maxcorr: the vectors were not resolved by the
// reduced-precision stage and a first-order step would only mix them), get
// the orthogonalising half alone.
__global__ void oa_fg(int n, const double* T, const double* P, const double* Q,
                      const double* sig, double guard2, double maxcorr, double* F, double* G,
                      int* rej) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / n), i = (int)(e - (int64_t)j * n);
    const double rij = (i == j ? 1.0 : 0.0) - P[e];
    const double qij = (i == j ? 1.0 : 0.0) - Q[e];
    if (i == j) {
      F[e] = 0.5 * rij;
      G[e] = 0.5 * qij;
      continue;
    }
    const double si = sig[i], sj = sig[j];
    const double d = (sj - si) * (sj + si);
    if (!(fabs(d) > guard2)) {
      F[e] = 0.5 * rij;
      G[e] = 0.5 * qij;
      rej[i] = 1;
      continue;
    }
    const double tij = T[e], tji = T[j + (int64_t)i * n];
    const double a = -tij - rij * sj;
    const double b = -tji - sj * qij;
    const double f = (-sj * a - si * b) / d, g = (-si * a - sj * b) / d;
    // the same pair seen from (j, i): the decision must agree on both sides
    // so that F + F^T = R and G + G^T = S hold for every pair
    const double a2 = -tji - rij * si;
    const double b2 = -tij - si * qij;
    const double f2 = (-si * a2 - sj * b2) / (-d), g2 = (-sj * a2 - si * b2) / (-d);
    const bool ok = fabs(f) <= maxcorr && fabs(g) <= maxcorr && fabs(f2) <= maxcorr &&
                    fabs(g2) <= maxcorr;
    F[e] = ok ? f : 0.5 * rij;
    G[e] = ok ? g : 0.5 * qij;
    // rej[i] = 1: index i belongs to a pair whose vectors the refinement
    // could not separate (the pair marks i here and j from (j, i))
    if (!ok) rej[i] = 1;
  }
}

// D (n x k, ld n) = the columns idx[0..k-1] of S (ld lds).
__global__ void mx_gather(int n, int k, const int* idx, const double* S, int lds, double* D) {
  const int64_t total = (int64_t)n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / n, i = e - c * n;
    D[e] = S[i + (int64_t)idx[c] * lds];
  }
}

// The columns idx[0..k-1] of D (ld ldd) = S (n x k, ld n); sd[idx[c]] = ss[c].
__global__ void mx_scatter(int n, int k, const int* idx, const double* S, double* D, int ldd,
                           const double* ss, double* sd) {
  const int64_t total = (int64_t)n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / n, i = e - c * n;
    D[i + (int64_t)idx[c] * ldd] = S[e];
    if (ss && i == 0) sd[idx[c]] = ss[c];
  }
}

struct MixedBufs {
  int lda_p, ldb;
  float* A32;
  double *Ap, *alpha, *beta, *sigma, *Vt, *Ut;
  double *T, *P, *Q, *W, *F, *G, *U2, *V2, *part;
  float *U32, *F32, *UF32;  // the update products U F, V G in FP32 (2 n^2 each)
};

constexpr int MX_PARTS = 1024;

// Leading dimension of the FP32 working copy: n rounded up to 16, plus 16
// (64 bytes) when that is a multiple of 512 (2 KiB), as for the FP64 path's
// copy.  At 8192^2 the shift took 9-10 ms off every mixed variant; a
// 128-byte shift added 33 ms.
static int mixed_lda(int n) {
  const int l = (int)round_up(n, 16);
  return (l % 512 == 0) ? l + 16 : l;
}

void mixed_plan(ArenaPlan& p, int n, int upd, bool resolve) {
  const int lda_p = mixed_lda(n);
  const int ldb = (int)round_up(n, 16);
  p.add<float>((size_t)lda_p * n);
  p.add<double>((size_t)lda_p * n);
  for (int i = 0; i < 3; ++i) p.add<double>(n);
  p.add<double>((size_t)ldb * n * 2);
  for (int i = 0; i < 8; ++i) p.add<double>((size_t)n * n);
  for (int i = 0; i < 3; ++i) p.add<float>((size_t)n * n * 2);
  p.add<double>(MX_PARTS);
  p.add<int>(n);  // members of unresolved pairs
  if (resolve) svd_device_plan(p, n, n);
  bidiag_plan_f32(p, n, n, lda_p, upd);
  dc_plan(p, n);
  twisted_plan(p, n);
  backtrans_plan_f32(p, n);
}

static double host_defect(int n, const double* M, double* part, hipStream_t st, int& rc) {
  hipLaunchKernelGGL(mx_defect, dim3(MX_PARTS), dim3(256), 0, st, n, M, part);
  std::vector<double> h(MX_PARTS);
  rc = DDC_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h.data(), part, sizeof(double) * MX_PARTS, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    rc = DDC_E_HIP;
    return 0.0;
  }
  double s = 0.0;
  for (double v : h) s += v;
  return sqrt(s);
}

// Resolve: the members of the pairs the last refinement step could not separate
// (their gaps lie below the FP32 stage's error; idx lists them, rej holds
// room for them on the device) get a Rayleigh-Ritz step in FP64 on their own
// left and right vectors,
//   A_R = U_R^T A V_R = P S Q^T,  U_R <- U_R P,  V_R <- V_R Q,  sigma_R <- S.
// The steps have already decoupled U_R, V_R from the other vectors, so the
// FP64 SVD of the block A_R separates the members; the values
// keep their ascending order because each block value lies within the
// block's own error of the value it replaces.  Temporaries: UR, VR, X, URn
// n x k; AR, Pb, Qb k x k; sb k.
static int mixed_resolve(int n, const double* A, int lda, double* U, int ldu, double* V, int ldv,
                         double* sigma, const std::vector<int>& idx, int* rej, double* UR,
                         double* VR, double* X, double* AR, double* Pb, double* Qb, double* URn,
                         double* sb, Arena& ws, hipStream_t st) {
  const int k = (int)idx.size();
  DDC_HIP(hipMemcpyAsync(rej, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
  const unsigned gk = (unsigned)std::min<int64_t>(ceil_div((int64_t)n * k, 256), 8192);
  hipLaunchKernelGGL(mx_gather, dim3(gk), dim3(256), 0, st, n, k, rej, U, ldu, UR);
  hipLaunchKernelGGL(mx_gather, dim3(gk), dim3(256), 0, st, n, k, rej, V, ldv, VR);
  DDC_LAUNCHED();
  DDC_TRY(dgemm('N', 'N', n, k, n, 1.0, A, lda, VR, n, 0.0, X, n, st));
  DDC_TRY(dgemm('T', 'N', k, k, n, 1.0, UR, n, X, n, 0.0, AR, k, st));
  DDC_TRY(svd_device(k, k, AR, k, sb, Pb, k, Qb, k, ws, st));
  DDC_TRY(dgemm('N', 'N', n, k, k, 1.0, UR, n, Pb, k, 0.0, URn, n, st));
  DDC_TRY(dgemm('N', 'N', n, k, k, 1.0, VR, n, Qb, k, 0.0, X, n, st));
  hipLaunchKernelGGL(mx_scatter, dim3(gk), dim3(256), 0, st, n, k, rej, URn, U, ldu, sb, sigma);
  hipLaunchKernelGGL(mx_scatter, dim3(gk), dim3(256), 0, st, n, k, rej, X, V, ldv, nullptr,
                     nullptr);
  DDC_LAUNCHED();
  return DDC_OK;
}

// diag (host, 8 entries, may be null): [0] ||I - U^T U||_F and [1]
// ||I - V^T V||_F before the first refinement step, [2] bidiagonalization,
// [3] values and vectors, [4] back-transformation, [5] refinement (resolve
// included), [6] the
// FP32 conversion, in milliseconds; [7] largest |change of sigma| / sigma_max
// in the last refinement step; with resolve (diag then holds 10 entries)
// [8] the number of values resolved and [9] the resolve stage's milliseconds.
int mixed_run(int n, const double* A, int lda, double* sigma_out, double* U, int ldu, double* V,
              int ldv, int upd, bool separated, bool resolve, int steps, double* diag,
              Arena& ws, hipStream_t st) {
  MixedBufs b;
  b.lda_p = mixed_lda(n);
  b.ldb = (int)round_up(n, 16);
  b.A32 = ws.take<float>((size_t)b.lda_p * n);
  b.Ap = ws.take<double>((size_t)b.lda_p * n);
  b.alpha = ws.take<double>(n);
  b.beta = ws.take<double>(n);
  b.sigma = ws.take<double>(n);
  b.Vt = ws.take<double>((size_t)b.ldb * n * 2);
  b.Ut = b.Vt ? b.Vt + (size_t)b.ldb * n : nullptr;
  double** mats[8] = {&b.T, &b.P, &b.Q, &b.W, &b.F, &b.G, &b.U2, &b.V2};
  for (auto pm : mats) *pm = ws.take<double>((size_t)n * n);
  b.U32 = ws.take<float>((size_t)n * n * 2);
  b.F32 = ws.take<float>((size_t)n * n * 2);
  b.UF32 = ws.take<float>((size_t)n * n * 2);
  b.part = ws.take<double>(MX_PARTS);
  int* rej = ws.take<int>(n);
  if (!rej) return DDC_E_NOMEM;

  hipEvent_t ev[8];
  for (auto& e : ev) DDC_HIP(hipEventCreate(&e));
  struct Ev {
    hipEvent_t* e;
    ~Ev() {
      for (int i = 0; i < 8; ++i) (void)hipEventDestroy(e[i]);
    }
  } evg{ev};
  DDC_HIP(hipEventRecord(ev[0], st));
  const int64_t tot = (int64_t)b.lda_p * n;
  const unsigned gconv = (unsigned)std::min<int64_t>(ceil_div(tot, 256), 8192);
  hipLaunchKernelGGL(mx_to_f32, dim3(gconv), dim3(256), 0, st, n, n, A, lda, b.A32, b.lda_p);
  DDC_LAUNCHED();
  DDC_HIP(hipEventRecord(ev[1], st));
  DDC_TRY(bidiag_run_f32(n, n, b.A32, b.lda_p, b.alpha, b.beta, upd, ws, st));
  DDC_HIP(hipEventRecord(ev[2], st));
  DDC_TRY(dc_run(n, b.alpha, b.beta, b.sigma, ws, st));
  DDC_TRY(twisted_run(n, b.alpha, b.beta, b.sigma, b.Vt, b.Ut, b.ldb, ws, st));
  DDC_HIP(hipEventRecord(ev[3], st));
  // back-transformation in FP32 (the refinement works from FP32-accurate U,
  // V either way: the reflectors are FP32), into the storage of Ap
  float* U32 = reinterpret_cast<float*>(b.Ap);
  float* V32 = U32 + (size_t)n * n;
  DDC_TRY(backtrans_run_f32(n, b.A32, b.lda_p, b.Ut, b.Vt, b.ldb, U32, V32, ws, st));
  double* U0 = steps > 0 ? b.U2 : U;
  double* V0 = steps > 0 ? b.V2 : V;
  const int ldu0 = steps > 0 ? n : ldu, ldv0 = steps > 0 ? n : ldv;
  const unsigned gsq = (unsigned)std::min<int64_t>(ceil_div((int64_t)n * n, 256), 8192);
  hipLaunchKernelGGL(mx_to_f64, dim3(gsq), dim3(256), 0, st, n, U32, U0, ldu0);
  hipLaunchKernelGGL(mx_to_f64, dim3(gsq), dim3(256), 0, st, n, V32, V0, ldv0);
  DDC_LAUNCHED();
  DDC_HIP(hipEventRecord(ev[4], st));

  double defect_u = 0.0, defect_v = 0.0, dsig = 0.0;
  std::vector<int> idx;  // the values resolved (host; read by the copy to rej)
  std::vector<double> s_old(n), s_new(n);
  double *Uc = b.U2, *Vc = b.V2, *X1 = b.W, *X2 = b.T;
  const dim3 gsy((unsigned)ceil_div(n, 32), (unsigned)ceil_div(n, 32));
  for (int it = 0; it < steps; ++it) {
    // current U, V in Uc, Vc (ld n); X1, X2 free
    DDC_TRY(dgemm('N', 'N', n, n, n, 1.0, A, lda, Vc, n, 0.0, X1, n, st));
    DDC_TRY(dgemm('T', 'N', n, n, n, 1.0, Uc, n, X1, n, 0.0, X2, n, st));
    DDC_TRY(dsyrk_upper_t(n, n, Uc, n, b.P, n, st));
    DDC_TRY(dsyrk_upper_t(n, n, Vc, n, b.Q, n, st));
    hipLaunchKernelGGL(mx_symm, gsy, dim3(32, 8), 0, st, n, b.P);
    hipLaunchKernelGGL(mx_symm, gsy, dim3(32, 8), 0, st, n, b.Q);
    DDC_LAUNCHED();
    if (it == 0 && diag) {
      int rc;
      defect_u = host_defect(n, b.P, b.part, st, rc);
      DDC_TRY(rc);
      defect_v = host_defect(n, b.Q, b.part, st, rc);
      DDC_TRY(rc);
    }
    DDC_HIP(hipMemcpyAsync(s_old.data(), b.sigma, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(oa_sigma, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n, X2, b.P,
                       b.Q, b.sigma);
    DDC_LAUNCHED();
    DDC_HIP(hipMemcpyAsync(s_new.data(), b.sigma, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    DDC_HIP(hipStreamSynchronize(st));
    double smax = 0.0;
    for (double v : s_new) smax = std::max(smax, fabs(v));
    dsig = 0.0;
    for (int i = 0; i < n; ++i) dsig = std::max(dsig, fabs(s_new[i] - s_old[i]));
    dsig = smax > 0 ? dsig / smax : 0.0;
    const double g = 1e-13 * smax * smax;
    const unsigned gfg = (unsigned)std::min<int64_t>(ceil_div((int64_t)n * n, 256), 16384);
    DDC_HIP(hipMemsetAsync(rej, 0, sizeof(int) * n, st));
    hipLaunchKernelGGL(oa_fg, dim3(gfg), dim3(256), 0, st, n, X2, b.P, b.Q, b.sigma, g,
                       separated ? OA_MAXCORR_SEPARATED : OA_MAXCORR, b.F, b.G, rej);
    DDC_LAUNCHED();
    // U <- U + U F, V <- V + V G into the free pair, or the outputs.  The
    // products are FP32 GEMMs: F and G are first-order corrections (entries
    // at most the correction bound), so the products' rounding, about
    // 6e-8 |U| |F|, is second order against the step's own error and
    // vanishes as F does.
    const bool last = it + 1 == steps;
    double* Un = last ? U : X1;
    double* Vn = last ? V : X2;
    const int ldun = last ? ldu : n, ldvn = last ? ldv : n;
    const size_t nn = (size_t)n * n;
    const unsigned gsq2 = (unsigned)std::min<int64_t>(ceil_div((int64_t)nn, 256), 8192);
    hipLaunchKernelGGL(mx_round_f32, dim3(gsq2), dim3(256), 0, st, n, Uc, b.U32);
    hipLaunchKernelGGL(mx_round_f32, dim3(gsq2), dim3(256), 0, st, n, Vc, b.U32 + nn);
    hipLaunchKernelGGL(mx_round_f32, dim3(gsq2), dim3(256), 0, st, n, b.F, b.F32);
    hipLaunchKernelGGL(mx_round_f32, dim3(gsq2), dim3(256), 0, st, n, b.G, b.F32 + nn);
    DDC_LAUNCHED();
    DDC_TRY(sgemm('N', 'N', n, n, n, 1.0f, b.U32, n, b.F32, n, 0.0f, b.UF32, n, st));
    DDC_TRY(sgemm('N', 'N', n, n, n, 1.0f, b.U32 + nn, n, b.F32 + nn, n, 0.0f, b.UF32 + nn, n,
                  st));
    hipLaunchKernelGGL(mx_add_f32, dim3(gsq2), dim3(256), 0, st, n, Uc, b.UF32, Un, ldun);
    hipLaunchKernelGGL(mx_add_f32, dim3(gsq2), dim3(256), 0, st, n, Vc, b.UF32 + nn, Vn, ldvn);
    DDC_LAUNCHED();
    if (!last) {
      std::swap(Uc, X1);
      std::swap(Vc, X2);
    }
    if (resolve && last) {
      // after the last step: the steps have decoupled the members from the
      // rest to second order by now (X1, X2 are free here).  Resolving after
      // the first step instead, so that later steps refine the members with
      // the rest, gave FP64-accurate values but a residual about 6x larger
      // (8192^2 graded: 1.2e-10 against 2.0e-11 with two steps).
      std::vector<int> hrej(n);
      DDC_HIP(hipMemcpyAsync(hrej.data(), rej, sizeof(int) * n, hipMemcpyDeviceToHost, st));
      DDC_HIP(hipStreamSynchronize(st));
      for (int i = 0; i < n; ++i)
        if (hrej[i]) idx.push_back(i);
      DDC_HIP(hipEventRecord(ev[6], st));
      if (!idx.empty())
        DDC_TRY(mixed_resolve(n, A, lda, Un, ldun, Vn, ldvn, b.sigma, idx, rej, X1, X2, b.F, b.P,
                              b.Q, b.G, reinterpret_cast<double*>(b.U32),
                              reinterpret_cast<double*>(b.F32), ws, st));
      DDC_HIP(hipEventRecord(ev[7], st));
    }
  }
  DDC_HIP(hipMemcpyAsync(sigma_out, b.sigma, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  DDC_HIP(hipEventRecord(ev[5], st));
  DDC_HIP(hipStreamSynchronize(st));
  if (diag) {
    float t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&t[i], ev[i], ev[i + 1]);
    if (resolve && steps > 0) (void)hipEventElapsedTime(&t[5], ev[6], ev[7]);
    diag[0] = defect_u;
    diag[1] = defect_v;
    diag[2] = t[1];
    diag[3] = t[2];
    diag[4] = t[3];
    diag[5] = t[4];
    diag[6] = t[0];
    diag[7] = dsig;
    if (resolve) {
      diag[8] = (double)idx.size();
      diag[9] = t[5];
    }
  }
  return DDC_OK;
}

}  // namespace ddc
