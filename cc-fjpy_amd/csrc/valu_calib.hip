// valu_calib.hip -- instruction-issue calibration kernels for gfx950
// (measurement only; not part of the transform library).
//
// Each kernel runs, per lane, `iters` loop iterations of 32 instructions of
// ONE class (8 independent register chains x 4), written as inline assembly
// so that the count is exact: 32 * iters wave64 instructions of that class
// per wave, plus the loop's 3 scalar instructions per iteration.  With the
// grid's wave count and the kernel time, this gives the issue cost of the
// class (cycles per wave64 instruction per SIMD) and, under rocprofv3 --pmc,
// what SQ_INSTS_VALU and its per-class counters read for a known count.
//
// A clock kernel reads the shader clock (s_memtime) and the 100 MHz
// reference counter (s_memrealtime) around a busy loop, for the engine clock
// the cycle counts need.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

enum Kind {
  K_FMA_F32, K_MUL_F32, K_ADD_F32, K_SIN_F32, K_RCP_F32, K_FMA_F64, K_FRACT_F64, K_ADD_U32,
  K_CNDMASK_B32, K_CVT_F32_F64, K_COUNT
};


template <int KIND>
__global__ __launch_bounds__(256) void k_calib(float* __restrict__ out, int iters, float seed) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[8];
  double d[8];
  uint32_t u[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    a[k] = seed + 1e-3f * k + 1e-7f * (float)threadIdx.x;
    d[k] = (double)a[k];
    u[k] = (uint32_t)tid + k;
  }
  const float x = 0.999f, y = 1e-4f;
  const double dx = 0.9999, dy = 1e-5;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (KIND == K_FMA_F32) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(x), "v"(y));
      if (KIND == K_MUL_F32) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a[k]) : "v"(x));
      if (KIND == K_ADD_F32) asm volatile("v_add_f32 %0, %0, %1" : "+v"(a[k]) : "v"(y));
      if (KIND == K_SIN_F32) asm volatile("v_sin_f32 %0, %0" : "+v"(a[k]));
      if (KIND == K_RCP_F32) asm volatile("v_rcp_f32 %0, %0" : "+v"(a[k]));
      if (KIND == K_FMA_F64) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[k]) : "v"(dx), "v"(dy));
      if (KIND == K_FRACT_F64) asm volatile("v_fract_f64 %0, %0" : "+v"(d[k]));
      if (KIND == K_ADD_U32) asm volatile("v_add_u32 %0, %0, %1" : "+v"(u[k]) : "v"((uint32_t)tid));
      if (KIND == K_CNDMASK_B32)
        asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(u[k]) : "v"((uint32_t)tid), "s"(mask));
      if (KIND == K_CVT_F32_F64) asm volatile("v_cvt_f32_f64 %0, %1" : "=v"(a[k]) : "v"(d[k]));
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) s += a[k] + (float)d[k] + (float)u[k];
  out[tid] = s;
}

__global__ void k_clock(unsigned long long* __restrict__ out, int spins) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long c0 = __builtin_amdgcn_s_memtime();
  const unsigned long long r0 = __builtin_amdgcn_s_memrealtime();
  float a = 1.0f;
  for (int i = 0; i < spins; i++) asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(a));
  const unsigned long long c1 = __builtin_amdgcn_s_memtime();
  const unsigned long long r1 = __builtin_amdgcn_s_memrealtime();
  out[0] = c1 - c0;
  out[1] = r1 - r0;
  out[2] = (unsigned long long)(a != a);
}

template <int KIND>
hipError_t launch(float* out, int nblocks, int iters, hipStream_t st) {
  hipLaunchKernelGGL((k_calib<KIND>), dim3(nblocks), dim3(256), 0, st, out, iters, 0.5f);
  return hipGetLastError();
}

}  // namespace

extern "C" {

// Number of instruction classes (kind 0 .. count-1):
// v_fma_f32, v_mul_f32, v_add_f32, v_sin_f32, v_rcp_f32, v_fma_f64,
// v_fract_f64, v_add_u32, v_cndmask_b32, v_cvt_f32_f64.
int calib_kinds(void) { return K_COUNT; }

// Runs class `kind` on nblocks x 256 lanes for `iters` iterations, out =
// device float[nblocks * 256]; *ms = kernel time (HIP events).  0 on success.
int calib_run(int kind, int nblocks, int iters, float* out, double* ms) {
  if (kind < 0 || kind >= K_COUNT || nblocks < 1 || iters < 1 || !out) return -1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -2;
  hipError_t e = hipEventRecord(e0, 0);
  switch (kind) {
    case K_FMA_F32: e = launch<K_FMA_F32>(out, nblocks, iters, 0); break;
    case K_MUL_F32: e = launch<K_MUL_F32>(out, nblocks, iters, 0); break;
    case K_ADD_F32: e = launch<K_ADD_F32>(out, nblocks, iters, 0); break;
    case K_SIN_F32: e = launch<K_SIN_F32>(out, nblocks, iters, 0); break;
    case K_RCP_F32: e = launch<K_RCP_F32>(out, nblocks, iters, 0); break;
    case K_FMA_F64: e = launch<K_FMA_F64>(out, nblocks, iters, 0); break;
    case K_FRACT_F64: e = launch<K_FRACT_F64>(out, nblocks, iters, 0); break;
    case K_ADD_U32: e = launch<K_ADD_U32>(out, nblocks, iters, 0); break;
    case K_CNDMASK_B32: e = launch<K_CNDMASK_B32>(out, nblocks, iters, 0); break;
    case K_CVT_F32_F64: e = launch<K_CVT_F32_F64>(out, nblocks, iters, 0); break;
  }
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return e == hipSuccess ? 0 : -3;
}

// Shader clock: out (device, 3 x uint64) = (s_memtime ticks, s_memrealtime
// ticks at 100 MHz, 0) over a busy loop of `spins` dependent FMAs.
int calib_clock(unsigned long long* out, int spins) {
  if (!out || spins < 1) return -1;
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, 0, out, spins);
  if (hipGetLastError() != hipSuccess) return -3;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -3;
}

}  // extern "C"
