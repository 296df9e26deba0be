// ccfj_internal.h -- helpers shared by the host-side sources of libccfj.so.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "../../include/ccfj.h"

namespace ccfj_internal {

// Record an error message for ccfj_last_error() and return code.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// Device allocation honouring the CCFJ_FAIL_ALLOC_AT fault injection.
int device_alloc(void** p, size_t bytes);
int ctx_device(const ccfj_ctx* ctx);
// Grow-only scratch buffer `which` (0 or 1) of the context, at least bytes.
int ctx_scratch(ccfj_ctx* ctx, int which, size_t bytes, void** p);
hipStream_t ctx_stream(const ccfj_ctx* ctx);
// With the context's timing on (ccfj_set_timing), events around a main
// kernel, summed by ccfj_get_timing; no-ops otherwise.
int timing_begin(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t>* ev);
void timing_end(ccfj_ctx* ctx, hipStream_t st, std::pair<hipEvent_t, hipEvent_t> ev);

}  // namespace ccfj_internal
