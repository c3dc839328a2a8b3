// capi.hip -- the C-ABI of libddcsvd (declared in include/svd_gpu.h).
//
// One process-wide context: the current HIP device at first use (so
// HIP_VISIBLE_DEVICES selects it; the reference asks on stdin instead,
// bidiag_par.c:151), one stream, and one grow-only device arena that every
// call carves its buffers from.  Calls are serialised by a mutex.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/svd_gpu.h"
#include "ddc_common.h"

namespace ddc {

namespace {
std::mutex g_mu;
char g_err[512] = "";
constexpr int kNumTimes = 12;
double g_times[kNumTimes] = {};

// Per-launch kernel clock (see kclock_on in ddc_common.h).
struct KClock {
  bool enabled = false;   // requested by svd_gpu_profile_kernels
  bool active = false;    // inside a timed call
  std::vector<hipEvent_t> pool;
  int used = 0;
  bool failed = false;
  std::vector<std::pair<int, int>> spans[KCLOCK_CLASSES];
} g_kc;
struct Ctx {
  bool init = false;
  int dev = 0;
  hipStream_t st = nullptr;
  char* arena = nullptr;
  size_t cap = 0;
  // pinned staging buffers of the host-buffer entry points, the events of
  // the copies that last used them and whether those may still be running
  char* stage[2] = {nullptr, nullptr};
  hipEvent_t sev[2] = {nullptr, nullptr};
  bool spend[2] = {false, false};
} g_ctx;

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace

bool kclock_on() { return g_kc.active && !g_kc.failed; }

int kclock_mark(hipStream_t st) {
  if (g_kc.used == (int)g_kc.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      g_kc.failed = true;
      return -1;
    }
    g_kc.pool.push_back(e);
  }
  if (hipEventRecord(g_kc.pool[g_kc.used], st) != hipSuccess) {
    g_kc.failed = true;
    return -1;
  }
  return g_kc.used++;
}

int kclock_pair(int cls, hipEvent_t* e0, hipEvent_t* e1) {
  for (int k = 0; k < 2; ++k) {
    if (g_kc.used + k == (int)g_kc.pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        g_kc.failed = true;
        return -1;
      }
      g_kc.pool.push_back(e);
    }
  }
  const int a = g_kc.used, b = g_kc.used + 1;
  g_kc.used += 2;
  *e0 = g_kc.pool[a];
  *e1 = g_kc.pool[b];
  kclock_span(cls, a, b);
  return 0;
}

void kclock_span(int cls, int a, int b) {
  if (a >= 0 && b >= 0 && cls >= 0 && cls < KCLOCK_CLASSES) g_kc.spans[cls].emplace_back(a, b);
}

static void kclock_begin() {
  g_kc.active = g_kc.enabled;
  g_kc.failed = false;
  g_kc.used = 0;
  for (auto& v : g_kc.spans) v.clear();
}

// The clock runs only inside one svd call, whatever path leaves it.
struct KClockScope {
  KClockScope() { kclock_begin(); }
  ~KClockScope() { g_kc.active = false; }
};

// After the stream has drained: milliseconds and launch counts per class
// into g_times[6..9] (left, right GEMV) and [10..11] (one-pass GEMV).
static void kclock_collect() {
  static const int slot_ms[KCLOCK_CLASSES] = {6, 7, 10}, slot_n[KCLOCK_CLASSES] = {8, 9, 11};
  for (int c = 0; c < KCLOCK_CLASSES; ++c) {
    double ms = 0.0;
    for (auto& pr : g_kc.spans[c]) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, g_kc.pool[pr.first], g_kc.pool[pr.second]) == hipSuccess) ms += t;
    }
    g_times[slot_ms[c]] = g_kc.active && !g_kc.failed ? ms : 0.0;
    g_times[slot_n[c]] = g_kc.active && !g_kc.failed ? (double)g_kc.spans[c].size() : 0.0;
  }
  g_kc.active = false;
}

void report_error(const char* msg) { set_err("%s", msg); }
void clear_error() { g_err[0] = 0; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_err("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return DDC_E_HIP;
}

static int ctx_init() {
  if (g_ctx.init) return DDC_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_err("no HIP device available");
    return DDC_E_NODEV;
  }
  DDC_HIP(hipGetDevice(&g_ctx.dev));
  DDC_HIP(hipStreamCreateWithFlags(&g_ctx.st, hipStreamNonBlocking));
  g_ctx.init = true;
  return DDC_OK;
}

static int arena_get(size_t bytes, Arena& ws, hipStream_t st) {
  if (bytes > g_ctx.cap) {
    if (g_ctx.arena) DDC_HIP(hipFree(g_ctx.arena));
    g_ctx.arena = nullptr;
    g_ctx.cap = 0;
    if (hipMalloc(&g_ctx.arena, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_err("device allocation of %zu bytes failed", bytes);
      return DDC_E_NOMEM;
    }
    g_ctx.cap = bytes;
  }
  ws.base = g_ctx.arena;
  ws.cap = g_ctx.cap;
  ws.used = 0;
  // test aid: with DDCSVD_POISON_ARENA set, every call starts from a
  // workspace of NaN bytes, so a buffer read before this call writes it
  // shows up in the results
  static const bool poison = getenv("DDCSVD_POISON_ARENA") != nullptr;
  if (poison) DDC_HIP(hipMemsetAsync(g_ctx.arena, 0xff, g_ctx.cap, st));
  return DDC_OK;
}

// ---- host <-> device copies of the host-buffer entry points ----------------
// A copy from or to pageable memory goes through two pinned staging buffers:
// the CPU copies one slice between the caller's buffer and a staging buffer
// while the DMA engine moves the other (a pageable hipMemcpy stages through
// the runtime's own buffers at a fraction of the link's rate).

constexpr size_t STAGE_BYTES = size_t(32) << 20;  // per staging buffer
constexpr size_t STAGE_MIN = size_t(4) << 20;     // smaller copies go directly

static int stage_init() {
  if (g_ctx.stage[1]) return DDC_OK;
  for (int b = 0; b < 2; ++b) {
    if (!g_ctx.sev[b]) DDC_HIP(hipEventCreateWithFlags(&g_ctx.sev[b], hipEventDisableTiming));
    if (!g_ctx.stage[b])
      DDC_HIP(hipHostMalloc(reinterpret_cast<void**>(&g_ctx.stage[b]), STAGE_BYTES,
                            hipHostMallocDefault));
  }
  return DDC_OK;
}

static int stage_wait(int b) {
  if (g_ctx.spend[b]) {
    DDC_HIP(hipEventSynchronize(g_ctx.sev[b]));
    g_ctx.spend[b] = false;
  }
  return DDC_OK;
}

// cols columns of width bytes: dst[c dpitch + i] = src[c spitch + i]; a
// large slice is split over four host threads
static void host_copy2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width,
                        size_t cols) {
  auto run = [=](size_t c0, size_t c1) {
    if (dpitch == width && spitch == width) {
      memcpy(dst + c0 * width, src + c0 * width, (c1 - c0) * width);
    } else {
      for (size_t c = c0; c < c1; ++c) memcpy(dst + c * dpitch, src + c * spitch, width);
    }
  };
  const bool contiguous = dpitch == width && spitch == width;
  const size_t bytes = width * cols;
  if (bytes < (size_t(8) << 20)) {
    run(0, cols);
    return;
  }
  constexpr int NT = 4;
  std::thread th[NT - 1];
  // parts [t * per, (t + 1) * per) of the bytes (contiguous) or of the
  // columns; a part whose thread cannot be started is copied here
  const size_t units = contiguous ? bytes : cols;
  const size_t per = (units + NT - 1) / NT;
  auto part = [=](int t) {
    const size_t u0 = std::min(units, t * per), u1 = std::min(units, (t + 1) * per);
    if (contiguous)
      memcpy(dst + u0, src + u0, u1 - u0);
    else
      run(u0, u1);
  };
  bool started[NT - 1] = {};
  for (int t = 1; t < NT; ++t) {
    try {
      th[t - 1] = std::thread(part, t);
      started[t - 1] = true;
    } catch (...) {
      started[t - 1] = false;
    }
  }
  part(0);
  for (int t = 1; t < NT; ++t) {
    if (started[t - 1])
      th[t - 1].join();
    else
      part(t);
  }
}

// host (pageable) -> device, stream-ordered on st; the caller synchronizes
static int copy_h2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                    size_t cols, hipStream_t st) {
  if (width * cols < STAGE_MIN || width > STAGE_BYTES) {
    DDC_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, cols, hipMemcpyHostToDevice, st));
    return DDC_OK;
  }
  DDC_TRY(stage_init());
  const size_t per = STAGE_BYTES / width;
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  for (size_t c0 = 0, i = 0; c0 < cols; ++i) {
    const size_t nc = std::min(per, cols - c0);
    const int b = (int)(i & 1);
    DDC_TRY(stage_wait(b));
    host_copy2d(g_ctx.stage[b], width, s + c0 * spitch, spitch, width, nc);
    DDC_HIP(hipMemcpy2DAsync(d + c0 * dpitch, dpitch, g_ctx.stage[b], width, width, nc,
                             hipMemcpyHostToDevice, st));
    DDC_HIP(hipEventRecord(g_ctx.sev[b], st));
    g_ctx.spend[b] = true;
    c0 += nc;
  }
  return DDC_OK;
}

// device -> host (pageable); complete when it returns
static int copy_d2h(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                    size_t cols, hipStream_t st) {
  if (width * cols < STAGE_MIN || width > STAGE_BYTES) {
    DDC_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, cols, hipMemcpyDeviceToHost, st));
    DDC_HIP(hipStreamSynchronize(st));
    return DDC_OK;
  }
  DDC_TRY(stage_init());
  const size_t per = STAGE_BYTES / width;
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  size_t pc0 = 0, pnc = 0;
  int pb = -1;
  for (size_t c0 = 0, i = 0; c0 < cols; ++i) {
    const size_t nc = std::min(per, cols - c0);
    const int b = (int)(i & 1);
    DDC_TRY(stage_wait(b));
    DDC_HIP(hipMemcpy2DAsync(g_ctx.stage[b], width, s + c0 * spitch, spitch, width, nc,
                             hipMemcpyDeviceToHost, st));
    DDC_HIP(hipEventRecord(g_ctx.sev[b], st));
    g_ctx.spend[b] = true;
    if (pb >= 0) {
      // the previous slice, while this one is in flight
      DDC_TRY(stage_wait(pb));
      host_copy2d(d + pc0 * dpitch, dpitch, g_ctx.stage[pb], width, width, pnc);
    }
    pc0 = c0;
    pnc = nc;
    pb = b;
    c0 += nc;
  }
  if (pb >= 0) {
    DDC_TRY(stage_wait(pb));
    host_copy2d(d + pc0 * dpitch, dpitch, g_ctx.stage[pb], width, width, pnc);
  }
  return DDC_OK;
}

// flag[0] = 1 if any entry of the m x n matrix is not finite; amax = the
// bits of max |A| (non-negative doubles order as their bit patterns), one
// atomic per wave
__global__ void k_check_finite(int m, int n, const double* A, int lda, int* flag,
                               unsigned long long* amax) {
  int64_t total = (int64_t)m * n;
  double mx = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = e / m, i = e - j * m;
    double v = A[i + j * lda];
    if (!isfinite(v)) flag[0] = 1;
    else mx = fmax(mx, fabs(v));
  }
  if (amax) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(amax, __double_as_longlong(mx));
  }
}

// A[i, j] = A[i, j] s1 s2 on an m x n matrix (two powers of two: a scale
// of 2^1029 for a matrix of subnormal entries is not itself a double)
__global__ void k_scale(int m, int n, double* A, int lda, double s1, double s2) {
  int64_t total = (int64_t)m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = e / m, i = e - j * m;
    A[i + j * lda] = A[i + j * lda] * s1 * s2;
  }
}

// Checks the m x n device matrix A for NaN and Inf (*nonfinite) and, when
// its largest entry lies outside [2^-450, 2^450], scales it in place by a
// power of two (exact) into [1/2, 1): *aexp is then the binary exponent
// that undoes it (x 2^aexp), else 0.  flag: 4 ints of device scratch.
static int check_and_scale(int m, int n, double* A, int lda, int* flag, hipStream_t st,
                           bool* nonfinite, int* aexp) {
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(flag + 2);
  DDC_HIP(hipMemsetAsync(flag, 0, 4 * sizeof(int), st));
  hipLaunchKernelGGL(k_check_finite, dim3(1024), dim3(256), 0, st, m, n, A, lda, flag, amax);
  DDC_LAUNCHED();
  int hflag[4] = {0, 0, 0, 0};
  DDC_HIP(hipMemcpyAsync(hflag, flag, sizeof(hflag), hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  *nonfinite = hflag[0] != 0;
  *aexp = 0;
  unsigned long long abits = 0;
  std::memcpy(&abits, hflag + 2, sizeof(abits));
  double amx = 0.0;
  std::memcpy(&amx, &abits, sizeof(amx));
  int e = 0;
  if (amx > 0.0) (void)std::frexp(amx, &e);
  if (!*nonfinite && amx > 0.0 && (e < -450 || e > 450)) {
    hipLaunchKernelGGL(k_scale, dim3(1024), dim3(256), 0, st, m, n, A, lda,
                       std::ldexp(1.0, -e / 2), std::ldexp(1.0, -e - (-e / 2)));
    DDC_LAUNCHED();
    *aexp = e;
  }
  return DDC_OK;
}

// v[0..len) *= 2^aexp on the device (undoes check_and_scale)
static int unscale(int len, double* v, int aexp, hipStream_t st) {
  if (aexp == 0 || len <= 0) return DDC_OK;
  hipLaunchKernelGGL(k_scale, dim3(1024), dim3(256), 0, st, len, 1, v, len,
                     std::ldexp(1.0, aexp / 2), std::ldexp(1.0, aexp - aexp / 2));
  DDC_LAUNCHED();
  return DDC_OK;
}

// Workspace of the tall (m >= n) pipeline.
struct TallBufs {
  int lda_p, ldb;
  double *Ap, *alpha, *beta, *sigma, *Vt, *Ut, *U, *V;
  int* flag;
};

// Leading dimension of the working copy of A: m rounded up to 16, plus 8
// (64 bytes) when that is a multiple of 256 (2 KiB).  Columns at a large
// power-of-two stride start at the same offset in every HBM channel
// interleave; the 64-byte shift took 8192 x 8192 from 699-701 ms to
// 691-694 ms (shifts of 256 and 512 bytes were slower than none).
static int tall_lda(int m) {
  const int l = (int)round_up(m, 16);
  return (l % 256 == 0) ? l + 8 : l;
}

static void tall_plan(ArenaPlan& p, int m, int n, bool want_uv_out) {
  int lda_p = tall_lda(m);
  int ldb = (int)round_up(n, 16);
  p.add<double>((size_t)lda_p * n);
  p.add<double>(n);
  p.add<double>(n);
  p.add<double>(n);
  p.add<double>((size_t)ldb * n * 2);
  if (want_uv_out) {
    p.add<double>((size_t)m * n);
    p.add<double>((size_t)n * n);
  }
  p.add<int>(64);
  bidiag_plan(p, m, n, lda_p);
  dc_plan(p, n);
  twisted_plan(p, n);
  backtrans_plan(p, m, n);
}

static int tall_take(Arena& ws, int m, int n, bool want_uv_out, TallBufs& b) {
  b.lda_p = tall_lda(m);
  b.ldb = (int)round_up(n, 16);
  b.Ap = ws.take<double>((size_t)b.lda_p * n);
  b.alpha = ws.take<double>(n);
  b.beta = ws.take<double>(n);
  b.sigma = ws.take<double>(n);
  b.Vt = ws.take<double>((size_t)b.ldb * n * 2);
  b.Ut = b.Vt ? b.Vt + (size_t)b.ldb * n : nullptr;
  b.U = b.V = nullptr;
  if (want_uv_out) {
    b.U = ws.take<double>((size_t)m * n);
    b.V = ws.take<double>((size_t)n * n);
  }
  b.flag = ws.take<int>(64);
  return b.flag ? DDC_OK : DDC_E_NOMEM;
}

struct Timer {
  hipEvent_t ev[7];
  int n = 0;
  hipStream_t st;
  int create(hipStream_t s) {
    st = s;
    for (int i = 0; i < 7; ++i) DDC_HIP(hipEventCreate(&ev[i]));
    return DDC_OK;
  }
  int mark() {
    if (n < 7) DDC_HIP(hipEventRecord(ev[n++], st));
    return DDC_OK;
  }
  void finish() {
    for (int i = 0; i < kNumTimes; ++i) g_times[i] = 0.0;
    kclock_collect();
    for (int i = 0; i + 1 < n && i < 6; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) g_times[i] = ms;
    }
  }
  ~Timer() {
    for (int i = 0; i < 7; ++i) (void)hipEventDestroy(ev[i]);
  }
};

// Runs the four stages on b.Ap (already holding A with zero padding).
// Results: sigma in b.sigma, U (m x n) in dU/ldu, V (n x n) in dV/ldv.
static int tall_stages(int m, int n, TallBufs& b, Arena& ws, double* dU, int ldu,
                       double* dV, int ldv, hipStream_t st, Timer& tm) {
  // A matrix whose largest entry lies outside [2^-450, 2^450] is scaled by
  // a power of two: the values and vectors stages work with sigma^2, which
  // would underflow or overflow.  sigma is scaled back at the end; U and V
  // do not change.
  bool nonfinite = false;
  int aexp = 0;
  DDC_TRY(check_and_scale(m, n, b.Ap, b.lda_p, b.flag, st, &nonfinite, &aexp));
  if (nonfinite) {
    set_err("input matrix holds NaN or Inf");
    return DDC_E_NUMERIC;
  }
  DDC_TRY(tm.mark());
  DDC_TRY(bidiag_run(m, n, b.Ap, b.lda_p, b.alpha, b.beta, ws, st));
  DDC_TRY(tm.mark());
  DDC_TRY(dc_run(n, b.alpha, b.beta, b.sigma, ws, st));
  DDC_TRY(tm.mark());
  DDC_TRY(twisted_run(n, b.alpha, b.beta, b.sigma, b.Vt, b.Ut, b.ldb, ws, st));
  DDC_TRY(tm.mark());
  DDC_TRY(backtrans_run(m, n, b.Ap, b.lda_p, b.Ut, b.Vt, b.ldb, dU, ldu, dV, ldv, ws, st));
  DDC_TRY(unscale(n, b.sigma, aexp, st));
  DDC_TRY(tm.mark());
  return DDC_OK;
}

// Workspace and body of an FP64 SVD of a device matrix (m >= n) inside
// another stage, on the caller's arena and stream (used by the mixed form
// for the block of singular values its refinement did not resolve).
void svd_device_plan(ArenaPlan& p, int m, int n) { tall_plan(p, m, n, false); }

int svd_device(int m, int n, const double* dA, int lda, double* dsigma, double* dU, int ldu,
               double* dV, int ldv, Arena& ws, hipStream_t st) {
  if (m < n || n < 1 || lda < m || ldu < m || ldv < n) return DDC_E_ARG;
  TallBufs b;
  DDC_TRY(tall_take(ws, m, n, false, b));
  Timer tm;
  DDC_TRY(tm.create(st));
  DDC_HIP(hipMemsetAsync(b.Ap, 0, sizeof(double) * (size_t)b.lda_p * n, st));
  DDC_HIP(hipMemcpy2DAsync(b.Ap, sizeof(double) * b.lda_p, dA, sizeof(double) * lda,
                           sizeof(double) * m, n, hipMemcpyDeviceToDevice, st));
  DDC_TRY(tall_stages(m, n, b, ws, dU, ldu, dV, ldv, st, tm));
  DDC_HIP(hipMemcpyAsync(dsigma, b.sigma, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  return DDC_OK;
}

static int check_dims(int m, int n) {
  if (m < 1 || n < 1) {
    set_err("invalid dimensions m=%d n=%d", m, n);
    return DDC_E_ARG;
  }
  if ((int64_t)m * n > ((int64_t)1 << 40)) {
    set_err("matrix too large");
    return DDC_E_ARG;
  }
  return DDC_OK;
}

// host buffers, the drop-in semantics
static int svd_host(int m, int n, double* A, double* sigma, double* U, double* V) {
  DDC_TRY(check_dims(m, n));
  if (!A || !sigma || !U || !V) {
    set_err("null pointer argument");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = g_ctx.st;
  const bool wide = m < n;
  const int mt = wide ? n : m, nt = wide ? m : n;  // tall problem dims
  // a wide A is factored through its transpose; its own bidiagonalization,
  // in the reference's m x (m+1) layout, is run afterwards on a padded copy
  // (Aw, leading dimension ldw) so that A is left holding the reflectors
  const int ldw = (int)round_up(m, 16);
  ArenaPlan plan;
  tall_plan(plan, mt, nt, true);
  if (wide) {
    plan.add<double>((size_t)ldw * n);
    plan.add<double>(2 * (size_t)n);
    bidiag_plan(plan, m, n, ldw);
  }
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  TallBufs b;
  DDC_TRY(tall_take(ws, mt, nt, true, b));
  double* Aw = wide ? ws.take<double>((size_t)ldw * n) : nullptr;
  double* wab = wide ? ws.take<double>(2 * (size_t)n) : nullptr;
  Timer tm;
  DDC_TRY(tm.create(st));
  KClockScope kcs;
  DDC_TRY(tm.mark());
  DDC_HIP(hipMemsetAsync(b.Ap, 0, sizeof(double) * (size_t)b.lda_p * nt, st));
  const size_t D = sizeof(double);
  if (!wide) {
    DDC_TRY(copy_h2d(b.Ap, D * b.lda_p, A, D * m, D * m, n, st));
  } else {
    DDC_HIP(hipMemsetAsync(Aw, 0, D * (size_t)ldw * n, st));
    DDC_TRY(copy_h2d(Aw, D * ldw, A, D * m, D * m, n, st));
    DDC_TRY(transpose_f64(m, n, Aw, ldw, b.Ap, b.lda_p, st));
  }
  DDC_TRY(tall_stages(mt, nt, b, ws, b.U, mt, b.V, nt, st, tm));
  if (wide) {
    // the reflectors left in A are dimensionless: an input of extreme
    // magnitude is scaled as in tall_stages (its alpha, beta are not kept)
    bool nonfinite = false;
    int aexp = 0;
    DDC_TRY(check_and_scale(m, n, Aw, ldw, b.flag, st, &nonfinite, &aexp));
    DDC_TRY(bidiag_run(m, n, Aw, ldw, wab, wab + n, ws, st));
  }
  // outputs
  DDC_TRY(copy_d2h(sigma, D * nt, b.sigma, D * nt, D * nt, 1, st));
  if (!wide) {
    DDC_TRY(copy_d2h(U, D * m, b.U, D * m, D * m, n, st));
    DDC_TRY(copy_d2h(V, D * n, b.V, D * n, D * n, n, st));
    // A is left holding the reflectors, as bidiag_par leaves it (bidiag_par.c:415-418)
    DDC_TRY(copy_d2h(A, D * m, b.Ap, D * b.lda_p, D * m, n, st));
  } else {
    // A^T = U' S V'^T  =>  A = V' S U'^T : U = V' (m x m), V = U' (n x m)
    DDC_TRY(copy_d2h(U, D * m, b.V, D * m, D * m, m, st));
    DDC_TRY(copy_d2h(V, D * n, b.U, D * n, D * n, m, st));
    DDC_TRY(copy_d2h(A, D * m, Aw, D * ldw, D * m, n, st));
  }
  DDC_TRY(tm.mark());
  DDC_HIP(hipStreamSynchronize(st));
  tm.finish();
  return DDC_OK;
}

// Counter-collection workload: a padded copy of dA, then `steps` left-GEMV
// (kind 1: right-GEMV, kind 2: one-pass GEMV) launches of its
// bidiagonalization and nothing else.
static int gemv_only_dev(int kind, int m, int n, const double* dA, int lda, int steps,
                         hipStream_t user) {
  DDC_TRY(check_dims(m, n));
  const bool right = kind == 1;
  if (kind == 2 && (!dA || m < n || n < 3 || m > 8192 || lda < m || steps < 1 || steps > n - 2)) {
    set_err("invalid arguments (need 8192 >= m >= n >= 3, lda >= m, 1 <= steps <= n-2)");
    return DDC_E_ARG;
  }
  if (kind == 0 && (!dA || m < n || n < 2 || lda < m || steps < 1 || steps > n - 1)) {
    set_err("invalid arguments (need m >= n >= 2, lda >= m, 1 <= steps <= n-1)");
    return DDC_E_ARG;
  }
  if (right && (!dA || m < n || n < 3 || lda < m || steps < 1 || steps > n - 2)) {
    set_err("invalid arguments (need m >= n >= 3, lda >= m, 1 <= steps <= n-2)");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = user;  // NULL: the null stream, ordered with the caller's other work
  const int lda_p = tall_lda(m);  // the layout of the SVD's working copy
  ArenaPlan plan;
  plan.add<double>((size_t)lda_p * n);
  plan.add<double>(n);
  bidiag_plan(plan, m, n, lda_p);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  double* Ap = ws.take<double>((size_t)lda_p * n);
  double* da = ws.take<double>(n);
  DDC_HIP(hipMemsetAsync(Ap, 0, sizeof(double) * (size_t)lda_p * n, st));
  DDC_HIP(hipMemcpy2DAsync(Ap, sizeof(double) * lda_p, dA, sizeof(double) * lda,
                           sizeof(double) * m, n, hipMemcpyDeviceToDevice, st));
  if (kind == 2)
    DDC_TRY(bidiag_fused_gemv_only(m, n, Ap, lda_p, da, steps, ws, st));
  else if (right)
    DDC_TRY(bidiag_right_gemv_only(m, n, Ap, lda_p, da, steps, ws, st));
  else
    DDC_TRY(bidiag_left_gemv_only(m, n, Ap, lda_p, da, steps, ws, st));
  DDC_HIP(hipStreamSynchronize(st));
  return DDC_OK;
}

static int svd_dev(int m, int n, const double* dA, int lda, double* dsigma, double* dU,
                   int ldu, double* dV, int ldv, hipStream_t user) {
  DDC_TRY(check_dims(m, n));
  const int mn = m < n ? m : n;
  if (!dA || !dsigma || !dU || !dV || lda < m || ldu < m || ldv < n) {
    set_err("invalid device arguments");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = user;  // NULL: the null stream, ordered with the caller's other work
  const bool wide = m < n;
  const int mt = wide ? n : m, nt = mn;
  ArenaPlan plan;
  tall_plan(plan, mt, nt, wide);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  TallBufs b;
  DDC_TRY(tall_take(ws, mt, nt, wide, b));
  Timer tm;
  DDC_TRY(tm.create(st));
  KClockScope kcs;
  DDC_TRY(tm.mark());
  DDC_HIP(hipMemsetAsync(b.Ap, 0, sizeof(double) * (size_t)b.lda_p * nt, st));
  if (!wide) {
    DDC_HIP(hipMemcpy2DAsync(b.Ap, sizeof(double) * b.lda_p, dA, sizeof(double) * lda,
                             sizeof(double) * m, n, hipMemcpyDeviceToDevice, st));
    DDC_TRY(tall_stages(mt, nt, b, ws, dU, ldu, dV, ldv, st, tm));
  } else {
    DDC_TRY(transpose_f64(m, n, dA, lda, b.Ap, b.lda_p, st));
    DDC_TRY(tall_stages(mt, nt, b, ws, b.U, mt, b.V, nt, st, tm));
    DDC_HIP(hipMemcpy2DAsync(dU, sizeof(double) * ldu, b.V, sizeof(double) * nt,
                             sizeof(double) * m, m, hipMemcpyDeviceToDevice, st));
    DDC_HIP(hipMemcpy2DAsync(dV, sizeof(double) * ldv, b.U, sizeof(double) * mt,
                             sizeof(double) * n, m, hipMemcpyDeviceToDevice, st));
  }
  DDC_HIP(hipMemcpyAsync(dsigma, b.sigma, sizeof(double) * nt, hipMemcpyDeviceToDevice, st));
  DDC_TRY(tm.mark());
  DDC_HIP(hipStreamSynchronize(st));
  tm.finish();
  return DDC_OK;
}

static int bidiag_host(int m, int n, double* A, double* alpha, double* beta) {
  DDC_TRY(check_dims(m, n));
  // beta has n-1 entries for m >= n and m entries for m < n (the N x (N+1)
  // upper bidiagonal of a wide matrix, bidiag.c:133-160)
  const int nbeta = m >= n ? n - 1 : m;
  if (!A || !alpha || (nbeta > 0 && !beta)) {
    set_err("null pointer argument");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = g_ctx.st;
  const int lda_p = (int)round_up(m, 16);
  ArenaPlan plan;
  plan.add<double>((size_t)lda_p * n);
  plan.add<double>(n);
  plan.add<double>(n);
  plan.add<int>(64);
  bidiag_plan(plan, m, n, lda_p);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  double* Ap = ws.take<double>((size_t)lda_p * n);
  double* da = ws.take<double>(n);
  double* db = ws.take<double>(n);
  int* flag = ws.take<int>(64);
  DDC_HIP(hipMemsetAsync(Ap, 0, sizeof(double) * (size_t)lda_p * n, st));
  DDC_HIP(hipMemcpy2DAsync(Ap, sizeof(double) * lda_p, A, sizeof(double) * m,
                           sizeof(double) * m, n, hipMemcpyHostToDevice, st));
  // extreme magnitudes: scaled by a power of two, alpha and beta scaled
  // back (the reflectors left in A are dimensionless); a matrix holding NaN
  // or Inf is bidiagonalized as it stands
  bool nonfinite = false;
  int aexp = 0;
  DDC_TRY(check_and_scale(m, n, Ap, lda_p, flag, st, &nonfinite, &aexp));
  DDC_TRY(bidiag_run(m, n, Ap, lda_p, da, db, ws, st));
  DDC_TRY(unscale(std::min(m, n), da, aexp, st));
  DDC_TRY(unscale(nbeta, db, aexp, st));
  DDC_HIP(hipMemcpy2DAsync(A, sizeof(double) * m, Ap, sizeof(double) * lda_p,
                           sizeof(double) * m, n, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipMemcpyAsync(alpha, da, sizeof(double) * std::min(m, n), hipMemcpyDeviceToHost, st));
  if (nbeta > 0)
    DDC_HIP(hipMemcpyAsync(beta, db, sizeof(double) * nbeta, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  return DDC_OK;
}

static int bdsvd_host(int n, const double* alpha, const double* beta, double* sigma,
                      int refine, double* X, double* Y) {
  if (n < 1 || !alpha || !sigma || (n > 1 && !beta)) {
    set_err("invalid arguments");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = g_ctx.st;
  const int ldb = (int)round_up(n, 16);
  const bool vec = refine || X;
  ArenaPlan plan;
  plan.add<double>(n);
  plan.add<double>(n);
  plan.add<double>(n);
  dc_plan(plan, n);
  if (vec) {
    plan.add<double>((size_t)ldb * n * 2);
    plan.add<double>((size_t)n * n);
    twisted_plan(plan, n);
  }
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  double* da = ws.take<double>(n);
  double* db = ws.take<double>(n);
  double* ds = ws.take<double>(n);
  // a bidiagonal whose largest entry lies outside [2^-450, 2^450] is scaled
  // by a power of two (exact, in two factors) and sigma scaled back, as in
  // the SVD: the values and vectors stages work with sigma^2
  double amx = 0.0;
  for (int i = 0; i < n; ++i) amx = std::fmax(amx, std::fabs(alpha[i]));
  for (int i = 0; i + 1 < n; ++i) amx = std::fmax(amx, std::fabs(beta[i]));
  int aexp = 0;  // stays 0 for a zero or non-finite bidiagonal
  if (amx > 0.0 && std::isfinite(amx)) (void)std::frexp(amx, &aexp);
  const bool scaled = aexp < -450 || aexp > 450;
  std::vector<double> sa, sb;
  const double* pa = alpha;
  const double* pb = beta;
  if (scaled) {
    const double f1 = std::ldexp(1.0, -aexp / 2), f2 = std::ldexp(1.0, -aexp - (-aexp / 2));
    sa.resize(n);
    for (int i = 0; i < n; ++i) sa[i] = alpha[i] * f1 * f2;
    pa = sa.data();
    if (n > 1) {
      sb.resize(n - 1);
      for (int i = 0; i + 1 < n; ++i) sb[i] = beta[i] * f1 * f2;
      pb = sb.data();
    }
  }
  DDC_HIP(hipMemcpyAsync(da, pa, sizeof(double) * n, hipMemcpyHostToDevice, st));
  DDC_HIP(hipMemsetAsync(db, 0, sizeof(double) * n, st));
  if (n > 1)
    DDC_HIP(hipMemcpyAsync(db, pb, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st));
  DDC_TRY(dc_run(n, da, db, ds, ws, st));
  if (vec) {
    double* Vt = ws.take<double>((size_t)ldb * n * 2);
    double* Ut = Vt + (size_t)ldb * n;
    double* tmp = ws.take<double>((size_t)n * n);
    DDC_TRY(twisted_run(n, da, db, ds, Vt, Ut, ldb, ws, st));
    if (X) {
      DDC_TRY(transpose_f64(n, n, Vt, ldb, tmp, n, st));
      DDC_HIP(hipMemcpyAsync(X, tmp, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, st));
      DDC_HIP(hipStreamSynchronize(st));
      DDC_TRY(transpose_f64(n, n, Ut, ldb, tmp, n, st));
      DDC_HIP(hipMemcpyAsync(Y, tmp, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, st));
    }
  }
  DDC_HIP(hipMemcpyAsync(sigma, ds, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  if (scaled) {
    const double g1 = std::ldexp(1.0, aexp / 2), g2 = std::ldexp(1.0, aexp - aexp / 2);
    for (int i = 0; i < n; ++i) sigma[i] = sigma[i] * g1 * g2;
  }
  return DDC_OK;
}

static int backward_dev(int m, int n, const double* U, int ldu, const double* sigma,
                        const double* V, int ldv, const double* gU, int ldgu,
                        const double* gs, const double* gV, int ldgv, double* gA, int ldga,
                        double guard, hipStream_t user) {
  DDC_TRY(check_dims(m, n));
  const int mn = m < n ? m : n;
  if (!U || !sigma || !V || !gA || ldu < m || ldv < n || ldga < m ||
      (gU && ldgu < m) || (gV && ldgv < n) || !(guard >= 0.0)) {
    set_err("invalid backward arguments");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = user;  // NULL: the null stream, ordered with the caller's other work
  const bool wide = m < n;
  const int mt = wide ? n : m;
  ArenaPlan plan;
  adjoint_plan(plan, mt, mn);
  if (wide) plan.add<double>((size_t)mt * mn);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  if (!wide) {
    DDC_TRY(adjoint_run(m, n, U, ldu, sigma, V, ldv, gU, ldgu, gs, gV, ldgv, gA, ldga, guard,
                        ws, st));
  } else {
    // A^T = V S U^T is tall: its adjoint is Abar^T
    double* tmp = ws.take<double>((size_t)mt * mn);
    if (!tmp) return DDC_E_NOMEM;
    DDC_TRY(adjoint_run(n, m, V, ldv, sigma, U, ldu, gV, ldgv, gs, gU, ldgu, tmp, n, guard,
                        ws, st));
    DDC_TRY(transpose_f64(n, m, tmp, n, gA, ldga, st));
  }
  DDC_HIP(hipStreamSynchronize(st));
  return DDC_OK;
}

// host buffers: A (m x n, reflectors as bidiag_gpu leaves them) -> Q
static int form_host(bool left, int m, int n, const double* A, double* Q) {
  DDC_TRY(check_dims(m, n));
  if (m < n) {
    set_err("bidiag_form_u_gpu / bidiag_form_v_gpu require m >= n");
    return DDC_E_ARG;
  }
  if (!A || !Q) {
    set_err("null pointer argument");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = g_ctx.st;
  const int lda_p = (int)round_up(m, 16);
  const int k = left ? m : n;
  ArenaPlan plan;
  plan.add<double>((size_t)lda_p * n);
  plan.add<double>((size_t)k * k);
  form_q_plan(plan, m, n, left);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  double* Ap = ws.take<double>((size_t)lda_p * n);
  double* Qd = ws.take<double>((size_t)k * k);
  if (!Qd) return DDC_E_NOMEM;
  DDC_HIP(hipMemsetAsync(Ap, 0, sizeof(double) * (size_t)lda_p * n, st));
  DDC_HIP(hipMemcpy2DAsync(Ap, sizeof(double) * lda_p, A, sizeof(double) * m,
                           sizeof(double) * m, n, hipMemcpyHostToDevice, st));
  DDC_TRY(form_q_run(left, m, n, Ap, lda_p, Qd, k, ws, st));
  DDC_HIP(hipMemcpyAsync(Q, Qd, sizeof(double) * (size_t)k * k, hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  return DDC_OK;
}

static int mixed_dev(int n, const double* dA, int lda, double* dsigma, double* dU, int ldu,
                     double* dV, int ldv, int update, int steps, double* diag, hipStream_t user) {
  DDC_TRY(check_dims(n, n));
  const int upd = update & 0xff;
  const int flags = update & ~0xff;
  if (!dA || !dsigma || !dU || !dV || lda < n || ldu < n || ldv < n || steps < 0 || steps > 8 ||
      (upd != UPD_F32 && upd != UPD_BF16X3 && upd != UPD_BF16) ||
      (flags & ~(SVD_GPU_MIXED_SEPARATED | SVD_GPU_MIXED_RESOLVE)) != 0 ||
      ((flags & SVD_GPU_MIXED_RESOLVE) != 0 && steps < 1)) {
    set_err("invalid svd_gpu_mixed arguments");
    return DDC_E_ARG;
  }
  DDC_TRY(ctx_init());
  hipStream_t st = user;  // NULL: the null stream, ordered with the caller's other work
  ArenaPlan plan;
  plan.add<int>(64);
  const bool resolve = (flags & SVD_GPU_MIXED_RESOLVE) != 0;
  mixed_plan(plan, n, upd, resolve);
  Arena ws;
  DDC_TRY(arena_get(plan.bytes, ws, st));
  int* flag = ws.take<int>(64);
  if (!flag) return DDC_E_NOMEM;
  DDC_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_check_finite, dim3(1024), dim3(256), 0, st, n, n, dA, lda, flag,
                     (unsigned long long*)nullptr);
  DDC_LAUNCHED();
  int hflag = 0;
  DDC_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  DDC_HIP(hipStreamSynchronize(st));
  if (hflag) {
    set_err("input matrix holds NaN or Inf");
    return DDC_E_NUMERIC;
  }
  DDC_TRY(mixed_run(n, dA, lda, dsigma, dU, ldu, dV, ldv, upd,
                    (flags & SVD_GPU_MIXED_SEPARATED) != 0, resolve, steps, diag, ws, st));
  return DDC_OK;
}

}  // namespace ddc

using namespace ddc;

extern "C" {

int svd_gpu_status(int m, int n, double* A, double* sigma, double* U, double* V) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return svd_host(m, n, A, sigma, U, V);
}

void svd_gpu(int m, int n, double* A, double* sigma, double* U, double* V) {
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_err[0] = 0;
    rc = svd_host(m, n, A, sigma, U, V);
  }
  if (rc != 0) {
    fprintf(stderr, "svd_gpu: %s (status %d)\n", g_err, rc);
    abort();
  }
}

int svd_gpu_dev(int m, int n, const double* dA, int lda, double* dsigma, double* dU,
                int ldu, double* dV, int ldv, void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return svd_dev(m, n, dA, lda, dsigma, dU, ldu, dV, ldv, (hipStream_t)stream);
}

int bidiag_gpu(int m, int n, double* A, double* alpha, double* beta) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return bidiag_host(m, n, A, alpha, beta);
}

int bdsvd_values_gpu(int n, const double* alpha, const double* beta, double* sigma,
                     int refine) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return bdsvd_host(n, alpha, beta, sigma, refine, nullptr, nullptr);
}

int bdsvd_vectors_gpu(int n, const double* alpha, const double* beta, double* sigma,
                      double* X, double* Y) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  if (!X || !Y) {
    set_err("null pointer argument");
    return DDC_E_ARG;
  }
  return bdsvd_host(n, alpha, beta, sigma, 1, X, Y);
}

int svd_gpu_backward(int m, int n, const double* U, int ldu, const double* sigma,
                     const double* V, int ldv, const double* gU, int ldgu,
                     const double* gsigma, const double* gV, int ldgv, double* gA, int ldga,
                     double gap_guard, void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return backward_dev(m, n, U, ldu, sigma, V, ldv, gU, ldgu, gsigma, gV, ldgv, gA, ldga,
                      gap_guard, (hipStream_t)stream);
}

int svd_gpu_mixed(int n, const double* dA, int lda, double* dsigma, double* dU, int ldu,
                  double* dV, int ldv, int update, int refine_steps, double* diag, void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return mixed_dev(n, dA, lda, dsigma, dU, ldu, dV, ldv, update, refine_steps, diag,
                   (hipStream_t)stream);
}

int bidiag_form_u_gpu(int m, int n, const double* A, double* U) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return form_host(true, m, n, A, U);
}

int bidiag_form_v_gpu(int m, int n, const double* A, double* V) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return form_host(false, m, n, A, V);
}

int svd_gpu_last_timings(double* ms, int count) {
  std::lock_guard<std::mutex> lk(g_mu);
  int k = count < kNumTimes ? count : kNumTimes;
  for (int i = 0; i < k; ++i) ms[i] = g_times[i];
  return k;
}

int svd_gpu_profile_left_gemv(int m, int n, const double* dA, int lda, int steps,
                              void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return gemv_only_dev(0, m, n, dA, lda, steps, (hipStream_t)stream);
}

int svd_gpu_profile_right_gemv(int m, int n, const double* dA, int lda, int steps,
                               void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return gemv_only_dev(1, m, n, dA, lda, steps, (hipStream_t)stream);
}

int svd_gpu_profile_fused_gemv(int m, int n, const double* dA, int lda, int steps,
                               void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  return gemv_only_dev(2, m, n, dA, lda, steps, (hipStream_t)stream);
}

int svd_gpu_profile_kernels(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_kc.enabled = on != 0;
  return DDC_OK;
}

int svd_gpu_dgemm(char transa, char transb, int M, int N, int K, double alpha, const double* dA,
                  int lda, const double* dB, int ldb, double beta, double* dC, int ldc,
                  void* stream) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err[0] = 0;
  const bool ta = transa == 'T' || transa == 't', tb = transb == 'T' || transb == 't';
  if (M < 0 || N < 0 || K < 0 || (!ta && transa != 'N' && transa != 'n') ||
      (!tb && transb != 'N' && transb != 'n') || lda < (ta ? K : M) || lda < 1 ||
      ldb < (tb ? N : K) || ldb < 1 || ldc < M || ldc < 1 || (M > 0 && N > 0 && !dC) ||
      (K > 0 && (!dA || !dB))) {
    set_err("invalid svd_gpu_dgemm arguments");
    return DDC_E_ARG;
  }
  DDC_TRY(dgemm_mfma(transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc,
                     (hipStream_t)stream));
  return DDC_OK;
}

const char* svd_gpu_last_error(void) { return g_err; }

}  // extern "C"
