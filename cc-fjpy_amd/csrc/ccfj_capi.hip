// ccfj_capi.hip -- host side of libccfj.so: contexts, workspace, the C ABI.
//
// Boundary (reference): the Cython layer hands float32 host buffers to
// FJ/FH (CPU build, include/FJcpu.h:19-20) or FJ/FHr/FHi (GPU build,
// include/FJgpu.hh:13-15).  The reference GPU host code allocates, copies,
// launches and frees on every call (src/FJgpu.cu:390-419).  Here a context
// keeps a stream and grow-only device buffers; the reference-compatible
// entry points use one default context per device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_kernels.h"
#include "fj_synth.h"

namespace ccfj_internal {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? CCFJ_ENOMEM : CCFJ_EDEVICE, "%s: %s (%d)",
              what, hipGetErrorString(e), (int)e);
}

}  // namespace ccfj_internal

namespace {

using ccfj_internal::fail;
using ccfj_internal::g_last_error;
using ccfj_internal::hip_fail;
using ccfj_internal::timing_begin;
using ccfj_internal::timing_end;

// Fault injection for tests: CCFJ_FAIL_ALLOC_AT=N makes the N-th device
// allocation of the process (1-based) fail with CCFJ_ENOMEM.
std::atomic<long> g_alloc_count{0};
bool injected_alloc_failure() {
  static long fail_at = [] {
    const char* s = getenv("CCFJ_FAIL_ALLOC_AT");
    return s ? atol(s) : 0L;
  }();
  long n = ++g_alloc_count;
  return fail_at > 0 && n == fail_at;
}

enum Slot { S_U0, S_U1, S_R, S_F, S_C, S_OUT, S_RR, S_HH, S_IR, S_ISR, S_BW, S_COEF, S_T, S_M, S_DR,
            S_X0,
            S_COUNT = S_X0 + ccfj_internal::kScratchSlots };

}  // namespace

struct ccfj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* buf[S_COUNT] = {};
  size_t cap[S_COUNT] = {};
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;  // recorded, unread
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  double timed_ms = 0.0;
  int timed_launches = 0;
  std::recursive_mutex mu;              // host-pointer entry points (ctx_lock)
  void* pin[2] = {nullptr, nullptr};   // pinned staging ring (ccfj_stage.hip)
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  std::mutex fft_mu;
  std::map<std::tuple<int, int, int>, hipfftHandle> fft_plans;  // (n, batch, type)
};

namespace {

int ensure(ccfj_ctx* ctx, Slot s, size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (ctx->cap[s] >= bytes) return CCFJ_OK;
  if (ctx->buf[s]) {
    // the buffer may still be read by work on any stream the caller used
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    (void)hipFree(ctx->buf[s]);
    ctx->buf[s] = nullptr;
    ctx->cap[s] = 0;
  }
  void* p = nullptr;
  int rc = ccfj_internal::device_alloc(&p, bytes);
  if (rc) return rc;
  ctx->buf[s] = p;
  ctx->cap[s] = bytes;
  return CCFJ_OK;
}

bool two_spectra(int mode) { return mode == CCFJ_MODE_J_MINUS_Y || mode == CCFJ_MODE_ABS_J || mode == CCFJ_MODE_ABS_H; }

int check_args(int mode, int itype, const void* U0, const void* U1, const void* r,
               const void* f, const void* c, int nr, int nf, int nc, const void* out) {
  if (mode < CCFJ_MODE_J || mode > CCFJ_MODE_ABS_H) return fail(CCFJ_EINVAL, "unknown mode %d", mode);
  if (itype != 0 && itype != 1) return fail(CCFJ_EINVAL, "itype must be 0 or 1, got %d", itype);
  if (nr < 0 || nf < 0 || nc < 0) return fail(CCFJ_EINVAL, "negative size (nr=%d nf=%d nc=%d)", nr, nf, nc);
  // one workgroup per 64 velocities of a frequency (the main kernel's tile)
  if ((long long)nf * ((nc + 63) / 64) > (1LL << 31) - 8) return fail(CCFJ_EINVAL, "grid too large");
  if (nr > 0 && nf > 0 && (!U0 || !r || !f)) return fail(CCFJ_EINVAL, "null input pointer");
  if (nr > 0 && nf > 0 && two_spectra(mode) && !U1) return fail(CCFJ_EINVAL, "mode %d needs U1", mode);
  if (nf > 0 && nc > 0 && (!c || !out)) return fail(CCFJ_EINVAL, "null c/out pointer");
  return CCFJ_OK;
}

int set_device(ccfj_ctx* ctx) {
  hipError_t e = hipSetDevice(ctx->device);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipSetDevice");
}

// Device work arrays of one transform (stations, records, results, column
// scale); rr..coef are unused when nr < 2, m when not normalising.
struct Work {
  double* rr = nullptr;
  float* hh = nullptr;
  float* ir = nullptr;
  float* isr = nullptr;
  uint32_t* bw = nullptr;
  float* coef = nullptr;
  float* T = nullptr;
  float* m = nullptr;
  float* dr = nullptr;
};
constexpr int kWork = 9;

// Bytes of each work array, in Work's order (256-byte multiples).
void work_sizes(int mode, int itype, int normalize, int nr, int nf, int nc, size_t sz[kWork]) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int nrp = fjk::padded(nr);
  const bool seg = nr >= 2;
  sz[0] = seg ? up(sizeof(double) * (nrp + 8)) : 0;
  sz[1] = sz[2] = sz[3] = sz[4] = seg ? up(sizeof(float) * (nrp + 8)) : 0;
  sz[5] = seg ? up(sizeof(float) * fjk::coef_floats(mode, itype, nrp, nf)) : 0;
  sz[6] = up(sizeof(float) * (size_t)nf * nc);
  sz[7] = normalize ? up(sizeof(float) * (size_t)nf) : 0;
  sz[8] = seg ? up(sizeof(float) * (nrp + 8)) : 0;
}

Work work_from(void* const ptr[kWork]) {
  Work w;
  w.rr = (double*)ptr[0];
  w.hh = (float*)ptr[1];
  w.ir = (float*)ptr[2];
  w.isr = (float*)ptr[3];
  w.bw = (uint32_t*)ptr[4];
  w.coef = (float*)ptr[5];
  w.T = (float*)ptr[6];
  w.m = (float*)ptr[7];
  w.dr = (float*)ptr[8];
  return w;
}

// Enqueue prep + main kernel (+ normalisation) on st.  Inputs and work
// arrays are device pointers; the main kernel is timed when `timed` and the
// context's timing is on.
int enqueue_transform(ccfj_ctx* ctx, int mode, int itype, int normalize, const float* U0,
                      const float* U1, const float* r, const float* f, const float* c, int nr,
                      int nf, int nc, float* out, hipStream_t st, const Work& w, bool timed) {
  if (nf == 0 || nc == 0) return CCFJ_OK;
  hipError_t e;
  int rc;
  if (nr < 2) {
    // no segment: the reference's sums are empty and stay 0 (even at f = 0,
    // where a segment would divide by k = 0)
    e = fjk::launch_fill(w.T, (size_t)nf * nc, 0.0f, st);
    if (e != hipSuccess) return hip_fail(e, "fill kernel launch");
  } else {
    fjk::Problem p;
    p.U0 = U0;
    p.U1 = two_spectra(mode) ? U1 : U0;
    p.r = r;
    p.f = f;
    p.c = c;
    p.nr = nr;
    p.nrp = fjk::padded(nr);
    p.nf = nf;
    p.nc = nc;
    p.rr = w.rr;
    p.hh = w.hh;
    p.ir = w.ir;
    p.isr = w.isr;
    p.bw = w.bw;
    p.dr = w.dr;
    p.coef = w.coef;
    p.T = w.T;
    e = fjk::launch_prep(mode, itype, p, st);
    if (e != hipSuccess) return hip_fail(e, "prep kernel launch");
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (timed && (rc = timing_begin(ctx, st, &ev))) return rc;
    e = fjk::launch_fj(mode, itype, p, st);
    if (e != hipSuccess) return hip_fail(e, "F-J kernel launch");
    if (timed) timing_end(ctx, st, ev);
  }
  e = fjk::launch_finish(w.T, normalize ? w.m : nullptr, 1, nf, nc, out, st);
  if (e != hipSuccess) return hip_fail(e, "finish kernel launch");
  return CCFJ_OK;
}

// The same with the work arrays taken from the context's grow-only slots.
int run_device(ccfj_ctx* ctx, int mode, int itype, int normalize, const float* U0,
               const float* U1, const float* r, const float* f, const float* c, int nr,
               int nf, int nc, float* out, hipStream_t st) {
  if (nf == 0 || nc == 0) return CCFJ_OK;
  size_t sz[kWork];
  work_sizes(mode, itype, normalize, nr, nf, nc, sz);
  const Slot slots[kWork] = {S_RR, S_HH, S_IR, S_ISR, S_BW, S_COEF, S_T, S_M, S_DR};
  void* ptr[kWork] = {};
  for (int i = 0; i < kWork; i++) {
    if (!sz[i]) continue;
    int rc = ensure(ctx, slots[i], sz[i]);
    if (rc) return rc;
    ptr[i] = ctx->buf[slots[i]];
  }
  const Work w = work_from(ptr);
  return enqueue_transform(ctx, mode, itype, normalize, U0, U1, r, f, c, nr, nf, nc, out, st, w,
                           true);
}

// ---- work arrays of ccfj_transform_dev_noalloc ----------------------------
// One array per GPU, held by that GPU from the library's load to the end of
// the process.  In host code `scratch` names a host-side stand-in: its
// device address on each GPU comes from hipGetSymbolAddress, resolved once
// per device when the first context on it is created, and only that
// address is handed to kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;      // per device; nullptr until resolved
std::vector<std::string> g_scratch_err;  // per device; why resolving failed

// With `device` current: resolve its copy of `scratch` unless done.
void resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) {
    g_scratch_ptr.resize(device + 1, nullptr);
    g_scratch_err.resize(device + 1);
  }
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch_ptr[device] = p;
}

int scratch_of(int device, void** p) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < (int)g_scratch_ptr.size() && g_scratch_ptr[device]) {
    *p = g_scratch_ptr[device];
    return CCFJ_OK;
  }
  const bool known = device < (int)g_scratch_err.size() && !g_scratch_err[device].empty();
  return fail(CCFJ_EDEVICE, "no-allocation work array of device %d unavailable: %s", device,
              known ? g_scratch_err[device].c_str() : "not resolved");
}

std::mutex g_default_mu;
std::vector<ccfj_ctx*> g_default;  // per device

int default_ctx(ccfj_ctx** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_default_mu);
  if ((int)g_default.size() <= dev) g_default.resize(dev + 1, nullptr);
  if (!g_default[dev]) {
    int rc = ccfj_create(dev, &g_default[dev]);
    if (rc) return rc;
  }
  *out = g_default[dev];
  return CCFJ_OK;
}

int ref_entry(int mode, float* u_f, float* r, float* f, float* out, float* c, int nc,
              int nr, int nf, int type) {
  ccfj_ctx* ctx;
  int rc = default_ctx(&ctx);
  if (rc) return rc;
  return ccfj_transform(ctx, mode, type, 0, u_f, nullptr, r, f, c, nr, nf, nc, out);
}

}  // namespace

namespace ccfj_internal {

int device_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (injected_alloc_failure())
    return fail(CCFJ_ENOMEM, "device allocation of %zu bytes failed (injected)", bytes);
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return CCFJ_OK;
}

int ctx_device(const ccfj_ctx* ctx) { return ctx->device; }

int ctx_scratch(ccfj_ctx* ctx, int which, size_t bytes, void** p) {
  if (which < 0 || which >= kScratchSlots) return fail(CCFJ_EINVAL, "scratch slot %d", which);
  const Slot s = (Slot)(S_X0 + which);
  int rc = ensure(ctx, s, bytes);
  if (rc) return rc;
  *p = ctx->buf[s];
  return CCFJ_OK;
}
hipStream_t ctx_stream(const ccfj_ctx* ctx) { return ctx->stream; }

int default_context(ccfj_ctx** out) { return default_ctx(out); }

std::unique_lock<std::recursive_mutex> ctx_lock(ccfj_ctx* ctx) {
  return std::unique_lock<std::recursive_mutex>(ctx->mu);
}

constexpr size_t kPinChunk = (size_t)32 << 20;

int ctx_pinned_ring(ccfj_ctx* ctx, void* buf[2], hipEvent_t ev[2], size_t* chunk) {
  for (int i = 0; i < 2; i++) {
    if (!ctx->pin[i]) {
      hipError_t e = hipHostMalloc(&ctx->pin[i], kPinChunk, hipHostMallocDefault);
      if (e != hipSuccess) {
        ctx->pin[i] = nullptr;
        return hip_fail(e, "hipHostMalloc");
      }
    }
    if (!ctx->pin_ev[i]) {
      hipError_t e = hipEventCreateWithFlags(&ctx->pin_ev[i], hipEventDisableTiming);
      if (e != hipSuccess) {
        ctx->pin_ev[i] = nullptr;
        return hip_fail(e, "hipEventCreate");
      }
    }
    buf[i] = ctx->pin[i];
    ev[i] = ctx->pin_ev[i];
  }
  *chunk = kPinChunk;
  return CCFJ_OK;
}

int cached_fft_plan(ccfj_ctx* ctx, int n, int batch, hipfftType type, hipStream_t st,
                    hipfftHandle* h) {
  std::lock_guard<std::mutex> lk(ctx->fft_mu);
  const auto key = std::make_tuple(n, batch, (int)type);
  auto it = ctx->fft_plans.find(key);
  if (it == ctx->fft_plans.end()) {
    hipfftHandle p = 0;
    int nn[1] = {n};
    const int nb = n / 2 + 1;
    const bool fwd = type == HIPFFT_D2Z || type == HIPFFT_R2C;
    const bool c2c = type == HIPFFT_Z2Z || type == HIPFFT_C2C;
    const hipfftResult r = hipfftPlanMany(&p, 1, nn, nullptr, 1, (fwd || c2c) ? n : nb, nullptr,
                                          1, (fwd && !c2c) ? nb : n, type, batch);
    if (r != HIPFFT_SUCCESS)
      return fail(CCFJ_EDEVICE, "hipfftPlanMany (n=%d, batch=%d, type=%d) failed", n, batch,
                  (int)type);
    it = ctx->fft_plans.emplace(key, p).first;
  }
  if (hipfftSetStream(it->second, st) != HIPFFT_SUCCESS)
    return fail(CCFJ_EDEVICE, "hipfftSetStream failed");
  *h = it->second;
  return CCFJ_OK;
}

int timing_begin(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t>* ev) {
  if (!ctx->timing) return CCFJ_OK;
  hipError_t e;
  if (!ctx->spare.empty()) {
    *ev = ctx->spare.back();
    ctx->spare.pop_back();
  } else {
    if ((e = hipEventCreate(&ev->first)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ev->second)) != hipSuccess) return hip_fail(e, "hipEventCreate");
  }
  (void)hipEventRecord(ev->first, st);
  return CCFJ_OK;
}

void timing_end(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t> ev) {
  if (!ctx->timing) return;
  (void)hipEventRecord(ev.second, st);
  ctx->events.push_back(ev);
}

}  // namespace ccfj_internal

extern "C" {

const char* ccfj_last_error(void) { return g_last_error.c_str(); }

int ccfj_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int ccfj_create(int device, ccfj_ctx** out) {
  if (!out) return fail(CCFJ_EINVAL, "null ctx pointer");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(CCFJ_EDEVICE, "no HIP GPU available");
  if (device < 0) {
    e = hipGetDevice(&device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  }
  if (device >= n) return fail(CCFJ_EINVAL, "device %d out of range (%d visible)", device, n);
  ccfj_ctx* ctx = new ccfj_ctx();
  ctx->device = device;
  int rc = set_device(ctx);
  if (rc) {
    delete ctx;
    return rc;
  }
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return hip_fail(e, "hipStreamCreate");
  }
  resolve_scratch(device);   // a failure is reported by the entry point that needs it
  *out = ctx;
  return CCFJ_OK;
}

int ccfj_destroy(ccfj_ctx* ctx) {
  if (!ctx) return CCFJ_OK;
  set_device(ctx);
  (void)hipStreamSynchronize(ctx->stream);
  // work of the device-pointer entry points may still run on callers' streams
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->fft_plans) (void)hipfftDestroy(kv.second);
  ctx->fft_plans.clear();
  for (int s = 0; s < S_COUNT; s++)
    if (ctx->buf[s]) (void)hipFree(ctx->buf[s]);
  for (auto& ev : ctx->events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  for (auto& ev : ctx->spare) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  for (int i = 0; i < 2; i++) {
    if (ctx->pin[i]) (void)hipHostFree(ctx->pin[i]);
    if (ctx->pin_ev[i]) (void)hipEventDestroy(ctx->pin_ev[i]);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CCFJ_OK;
}

int ccfj_transform_dev(ccfj_ctx* ctx, int mode, int itype, int normalize, const float* U0,
                       const float* U1, const float* r, const float* f, const float* c,
                       int nr, int nf, int nc, float* out, void* stream) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = check_args(mode, itype, U0, U1, r, f, c, nr, nf, nc, out);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return run_device(ctx, mode, itype, normalize, U0, U1, r, f, c, nr, nf, nc, out, st);
}

int ccfj_transform_dev_noalloc(ccfj_ctx* ctx, int mode, int itype, int normalize,
                               const float* U0, const float* U1, const float* r, const float* f,
                               const float* c, int nr, int nf, int nc, float* out,
                               void* stream) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = check_args(mode, itype, U0, U1, r, f, c, nr, nf, nc, out);
  if (rc) return rc;
  if (nf == 0 || nc == 0) return CCFJ_OK;
  size_t sz[kWork], total = 0;
  work_sizes(mode, itype, normalize, nr, nf, nc, sz);
  for (int i = 0; i < kWork; i++) total += sz[i];
  if (total > kNoAllocScratchBytes)
    return fail(CCFJ_EINVAL,
                "ccfj_transform_dev_noalloc: nr=%d nf=%d nc=%d needs %zu bytes of work arrays, "
                "more than the %zu bytes it holds per GPU (use ccfj_transform_dev)",
                nr, nf, nc, total, kNoAllocScratchBytes);
  if ((rc = set_device(ctx))) return rc;
  void* base = nullptr;
  if ((rc = scratch_of(ctx->device, &base))) return rc;
  char* q = (char*)base;
  void* ptr[kWork] = {};
  for (int i = 0; i < kWork; i++) {
    if (sz[i]) ptr[i] = q;
    q += sz[i];
  }
  const Work w = work_from(ptr);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return enqueue_transform(ctx, mode, itype, normalize, U0, U1, r, f, c, nr, nf, nc, out, st, w,
                           false);
}

int ccfj_transform(ccfj_ctx* ctx, int mode, int itype, int normalize, const float* U0,
                   const float* U1, const float* r, const float* f, const float* c, int nr,
                   int nf, int nc, float* out) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  auto lk = ccfj_internal::ctx_lock(ctx);
  int rc = check_args(mode, itype, U0, U1, r, f, c, nr, nf, nc, out);
  if (rc) return rc;
  if (nf == 0 || nc == 0) return CCFJ_OK;
  if ((rc = set_device(ctx))) return rc;
  const size_t nu = (size_t)nr * nf;
  const bool two = two_spectra(mode);
  if ((rc = ensure(ctx, S_U0, sizeof(float) * nu))) return rc;
  if (two && (rc = ensure(ctx, S_U1, sizeof(float) * nu))) return rc;
  if ((rc = ensure(ctx, S_R, sizeof(float) * nr))) return rc;
  if ((rc = ensure(ctx, S_F, sizeof(float) * nf))) return rc;
  if ((rc = ensure(ctx, S_C, sizeof(float) * nc))) return rc;
  if ((rc = ensure(ctx, S_OUT, sizeof(float) * (size_t)nf * nc))) return rc;
  hipStream_t st = ctx->stream;
  hipError_t e = hipSuccess;
  if (nu) {
    if ((rc = ccfj_internal::copy_h2d(ctx, ctx->buf[S_U0], U0, sizeof(float) * nu, st))) return rc;
    if (two && (rc = ccfj_internal::copy_h2d(ctx, ctx->buf[S_U1], U1, sizeof(float) * nu, st)))
      return rc;
    e = hipMemcpyAsync(ctx->buf[S_R], r, sizeof(float) * nr, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ctx->buf[S_F], f, sizeof(float) * nf, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->buf[S_C], c, sizeof(float) * nc, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  rc = run_device(ctx, mode, itype, normalize, (const float*)ctx->buf[S_U0],
                  (const float*)ctx->buf[S_U1], (const float*)ctx->buf[S_R],
                  (const float*)ctx->buf[S_F], (const float*)ctx->buf[S_C], nr, nf, nc,
                  (float*)ctx->buf[S_OUT], st);
  if (rc) return rc;
  return ccfj_internal::copy_d2h(ctx, out, ctx->buf[S_OUT], sizeof(float) * (size_t)nf * nc, st);
}

int ccfj_transform_items(ccfj_ctx* ctx, int mode, int itype, const float* U0, const float* U1,
                         const float* r, const float* f, const float* c, int nr, int nf, int nc,
                         const int* items, int nitems, float* out, void* stream) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  if (mode != CCFJ_MODE_J && mode != CCFJ_MODE_J_MINUS_Y)
    return fail(CCFJ_EINVAL, "item lists support modes J and J_MINUS_Y, got %d", mode);
  int rc = check_args(mode, itype, U0, U1, r, f, c, nr, nf, nc, out);
  if (rc) return rc;
  if (nitems < 0 || (nitems > 0 && !items)) return fail(CCFJ_EINVAL, "bad item list");
  if (nitems == 0) return CCFJ_OK;
  if ((rc = set_device(ctx))) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (nr < 2) {   // no segment: listed entries are 0, as in the full grid
    hipError_t e = fjk::launch_fj_items(mode, itype, fjl::Stations{}, nullptr, f, c, items,
                                        nitems, nr, 0, nf, nc, out, st);
    return e == hipSuccess ? CCFJ_OK : hip_fail(e, "item kernel launch");
  }
  const int nrp = fjk::padded(nr);
  if ((rc = ensure(ctx, S_RR, sizeof(double) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_HH, sizeof(float) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_IR, sizeof(float) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_ISR, sizeof(float) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_BW, sizeof(uint32_t) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_DR, sizeof(float) * (nrp + 8)))) return rc;
  if ((rc = ensure(ctx, S_COEF, sizeof(float) * fjk::coef_floats(mode, itype, nrp, nf)))) return rc;
  fjk::Problem p;
  p.U0 = U0;
  p.U1 = mode == CCFJ_MODE_J_MINUS_Y ? U1 : U0;
  p.r = r;
  p.f = f;
  p.c = c;
  p.nr = nr;
  p.nrp = nrp;
  p.nf = nf;
  p.nc = nc;
  p.rr = (double*)ctx->buf[S_RR];
  p.hh = (float*)ctx->buf[S_HH];
  p.ir = (float*)ctx->buf[S_IR];
  p.isr = (float*)ctx->buf[S_ISR];
  p.bw = (uint32_t*)ctx->buf[S_BW];
  p.dr = (float*)ctx->buf[S_DR];
  p.coef = (float*)ctx->buf[S_COEF];
  p.T = nullptr;   // the item kernel writes out[c][f] directly
  hipError_t e = fjk::launch_prep(mode, itype, p, st);
  if (e != hipSuccess) return hip_fail(e, "prep kernel launch");
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if ((rc = timing_begin(ctx, st, &ev))) return rc;
  e = fjk::launch_fj_items(mode, itype, fjk::stations(p), p.coef, f, c, items, nitems, nr, nrp, nf,
                           nc, out, st);
  if (e != hipSuccess) return hip_fail(e, "item kernel launch");
  timing_end(ctx, st, ev);
  return CCFJ_OK;
}

int ccfj_synth_fields(ccfj_ctx* ctx, int B, int nr, int nf, const float* r, const float* f,
                      const float* modes, const float* sigma, unsigned long long seed, float* U,
                      void* stream) {
  if (!ctx || !r || !f || !modes || !sigma || !U) return fail(CCFJ_EINVAL, "null argument");
  if (B < 1 || nr < 1 || nf < 1) return fail(CCFJ_EINVAL, "bad sizes");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipError_t e = fjk::launch_synth_fields(r, f, modes, sigma, B, nr, nf, seed, U, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "synth fields launch");
}

int ccfj_synth_labels(ccfj_ctx* ctx, int B, int H, int nf, const float* cgrid, const float* f,
                      const float* modes, float halfwidth, float amin, float* Y, void* stream) {
  if (!ctx || !cgrid || !f || !modes || !Y) return fail(CCFJ_EINVAL, "null argument");
  if (B < 1 || H < 1 || nf < 1) return fail(CCFJ_EINVAL, "bad sizes");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipError_t e = fjk::launch_synth_labels(cgrid, f, modes, B, H, nf, halfwidth, amin, Y, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "synth labels launch");
}

int ccfj_pool_abs_max(ccfj_ctx* ctx, const float* in, int B, int H, int S, int W, float* out,
                      void* stream) {
  if (!ctx || !in || !out) return fail(CCFJ_EINVAL, "null argument");
  if (B < 1 || H < 1 || S < 1 || W < 1) return fail(CCFJ_EINVAL, "bad sizes");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipError_t e = fjk::launch_pool_abs_max(in, B, H, S, W, out, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "pool kernel launch");
}

int ccfj_pool_abs_max_ranges(ccfj_ctx* ctx, const float* in, int B, int N, int H, int W,
                             const int* ranges, float* out, void* stream) {
  if (!ctx || !in || !ranges || !out) return fail(CCFJ_EINVAL, "null argument");
  if (B < 1 || N < 1 || H < 1 || W < 1) return fail(CCFJ_EINVAL, "bad sizes");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipError_t e = fjk::launch_pool_abs_max_ranges(in, B, N, H, W, ranges, out, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "pool kernel launch");
}

int ccfj_label_argmax(ccfj_ctx* ctx, const float* X, int B, int H, int W, float* Y,
                      void* stream) {
  if (!ctx || !X || !Y) return fail(CCFJ_EINVAL, "null argument");
  if (B < 1 || H < 1 || W < 1) return fail(CCFJ_EINVAL, "bad sizes");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipError_t e = fjk::launch_label_argmax(X, B, H, W, Y, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "label kernel launch");
}

int ccfj_malloc(ccfj_ctx* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return fail(CCFJ_EINVAL, "null argument");
  int rc = set_device(ctx);
  if (rc) return rc;
  return ccfj_internal::device_alloc(ptr, bytes);
}

int ccfj_free(ccfj_ctx* ctx, void* ptr) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && ptr) e = hipFree(ptr);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipFree");
}

int ccfj_memcpy(ccfj_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
  if (!ctx || (bytes && (!dst || !src))) return fail(CCFJ_EINVAL, "null argument");
  if (kind < 0 || kind > 2) return fail(CCFJ_EINVAL, "bad copy kind %d", kind);
  int rc = set_device(ctx);
  if (rc) return rc;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                        : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, k, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipMemcpyAsync");
}

int ccfj_synchronize(ccfj_ctx* ctx) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipStreamSynchronize");
}

int ccfj_set_timing(ccfj_ctx* ctx, int enable) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  ctx->timing = enable != 0;
  return CCFJ_OK;
}

int ccfj_get_timing(ccfj_ctx* ctx, double* kernel_ms, int* launches) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = set_device(ctx);
  if (rc) return rc;
  for (auto& ev : ctx->events) {
    hipError_t e = hipEventSynchronize(ev.second);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, ev.first, ev.second);
    if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    ctx->timed_ms += ms;
    ctx->timed_launches++;
    ctx->spare.push_back(ev);
  }
  ctx->events.clear();
  if (kernel_ms) *kernel_ms = ctx->timed_ms;
  if (launches) *launches = ctx->timed_launches;
  ctx->timed_ms = 0.0;
  ctx->timed_launches = 0;
  return CCFJ_OK;
}

int FJ(float* u_f, float* r, float* f, float* out, float* c, int nc, int nr, int nf,
       int type) {
  return ref_entry(CCFJ_MODE_J, u_f, r, f, out, c, nc, nr, nf, type);
}

int FHr(float* u_f, float* r, float* f, float* out, float* c, int nc, int nr, int nf,
        int type) {
  return ref_entry(CCFJ_MODE_J, u_f, r, f, out, c, nc, nr, nf, type);
}

int FHi(float* u_f, float* r, float* f, float* out, float* c, int nc, int nr, int nf,
        int type) {
  return ref_entry(CCFJ_MODE_Y, u_f, r, f, out, c, nc, nr, nf, type);
}

int FH(float* u_f, float* r, float* f, float* out, float* c, int nc, int nr, int nf,
       int type, int nThread) {
  (void)nThread;
  return ref_entry(CCFJ_MODE_Y, u_f, r, f, out, c, nc, nr, nf, type);
}

}  // extern "C"
