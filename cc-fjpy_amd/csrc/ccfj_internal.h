// ccfj_internal.h -- helpers shared by the host-side sources of libccfj.so.
#pragma once
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <mutex>
#include <utility>

#include "../../include/ccfj.h"

namespace ccfj_internal {

// Record an error message for ccfj_last_error() and return code.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// Device allocation honouring the CCFJ_FAIL_ALLOC_AT fault injection.
int device_alloc(void** p, size_t bytes);
int ctx_device(const ccfj_ctx* ctx);
// Grow-only scratch buffer `which` (0 .. kScratchSlots-1) of the context,
// at least bytes; 0 and 1 belong to the Hilbert transform.
constexpr int kScratchSlots = 10;
int ctx_scratch(ccfj_ctx* ctx, int which, size_t bytes, void** p);
hipStream_t ctx_stream(const ccfj_ctx* ctx);
// The process-wide context of the current device (the reference-compatible
// entry points use it).
int default_context(ccfj_ctx** out);
// A hipFFT plan for batch packed rows of length n (R2C/D2Z forward,
// C2R/Z2D backward, C2C/Z2Z either way), bound to stream st.  Plans are
// cached in the context (a plan's work area must not be shared by two
// contexts, whose calls may run at once on different streams) and
// destroyed with it.
int cached_fft_plan(ccfj_ctx* ctx, int n, int batch, hipfftType type, hipStream_t st,
                    hipfftHandle* h);
// With the context's timing on (ccfj_set_timing), events around a main
// kernel, summed by ccfj_get_timing; no-ops otherwise.
int timing_begin(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t>* ev);
void timing_end(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t> ev);

// Held by the host-pointer (synchronous) entry points for the whole call,
// so threads sharing a context (the Python API's per-device default) run
// them one at a time, as the reference's GIL-holding calls did.
std::unique_lock<std::recursive_mutex> ctx_lock(ccfj_ctx* ctx);
// The context's pinned staging ring: two host buffers of *chunk bytes and
// an event for each (allocated on first use, freed with the context).
int ctx_pinned_ring(ccfj_ctx* ctx, void* buf[2], hipEvent_t ev[2], size_t* chunk);
// Synchronous copies for host entry points: large ones go through the
// pinned ring with the host side split over worker threads.
int copy_d2h(ccfj_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st);
int copy_h2d(ccfj_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st);

}  // namespace ccfj_internal
