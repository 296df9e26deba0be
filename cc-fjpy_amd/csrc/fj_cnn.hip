// fj_cnn.hip -- band-proposal CNN for the coarse-to-fine F-J: forward and
// backward kernels, loss and optimizer.
//
// No reference counterpart.  The network maps a coarse F-J spectrum
// (1 x H x W, H velocities x W frequencies, normalised per column) to
// per-(velocity, frequency) logits of "a dispersion ridge passes here":
//   conv3x3(1->32)+ReLU, conv3x3(32->32)+ReLU, conv3x3(32->32)+ReLU,
//   conv3x3(32->1)
// Weights are stored in bfloat16 (an fp32 master copy is kept for the
// optimizer), every product is accumulated in fp32.  Tensors are NCHW fp32.
//
// Kernels (one thread per output pixel of a 16x16 tile, the input tile with
// its halo staged in LDS, weights wave-uniform so they arrive by scalar
// loads):
//   k_conv<CI, CO, BWD>  BWD = false: y = conv(x, w) + b (optional ReLU)
//                        BWD = true : dx = conv^T(dy * mask, w) -- the
//                        gradient w.r.t. the input; mask = (y_prev > 0)
//   k_conv_dw<CI, CO>    dW = sum over pixels of (dy * mask) x-patch
//   k_bias_grad_part/final  dB = sum over pixels of (dy * mask), two passes
//   k_bce                logits -> loss sum and dL/dlogit (mean BCE)
//   k_adam               Adam on the fp32 master, bf16 copy rewritten
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_cnn.h"

namespace fjk {

__device__ inline float bf16_to_f32(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

__device__ inline uint16_t f32_to_bf16(float f) {
  // round to nearest even; NaN kept quiet
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int TH = 16, TW = 16;

// Weight of the forward layer w[co][ci][kh][kw]: the bf16 weights widened
// to fp32 (exactly their bf16 values; the kernels then take them as
// wave-uniform fp32 operands with no conversion per use).  In the backward
// (input-gradient) pass the roles swap: output channel = forward ci, input
// channel = forward co, taps flipped.
template <int CI, int CO, bool BWD>
__global__ __launch_bounds__(256) void k_conv(const float* __restrict__ x,
                                             const float* __restrict__ mask,
                                             const float* __restrict__ w,
                                             const float* __restrict__ bias, int H, int W,
                                             int relu, float* __restrict__ y) {
  __shared__ float xs[CI][TH + 2][TW + 2];
  const int b = blockIdx.z;
  const int h0 = blockIdx.y * TH, w0 = blockIdx.x * TW;
  const int ty = threadIdx.x / TW, tx = threadIdx.x % TW;
  const size_t plane = (size_t)H * W;
  const float* xb = x + (size_t)b * CI * plane;
  const float* mb = mask ? mask + (size_t)b * CI * plane : nullptr;
  for (int e = threadIdx.x; e < CI * (TH + 2) * (TW + 2); e += 256) {
    const int ci = e / ((TH + 2) * (TW + 2));
    const int rem = e - ci * (TH + 2) * (TW + 2);
    const int yy = rem / (TW + 2), xx = rem - yy * (TW + 2);
    const int hh = h0 + yy - 1, ww = w0 + xx - 1;
    float v = 0.0f;
    if (hh >= 0 && hh < H && ww >= 0 && ww < W) {
      const size_t idx = (size_t)ci * plane + (size_t)hh * W + ww;
      v = xb[idx];
      if (mb && !(mb[idx] > 0.0f)) v = 0.0f;
    }
    xs[ci][yy][xx] = v;
  }
  __syncthreads();
  const int h = h0 + ty, wc = w0 + tx;
  float acc[CO];
#pragma unroll
  for (int co = 0; co < CO; co++) acc[co] = (!BWD && bias) ? bias[co] : 0.0f;
#pragma unroll 2
  for (int ci = 0; ci < CI; ci++) {
#pragma unroll
    for (int kh = 0; kh < 3; kh++) {
#pragma unroll
      for (int kw = 0; kw < 3; kw++) {
        const float xv = xs[ci][ty + kh][tx + kw];
#pragma unroll
        for (int co = 0; co < CO; co++) {
          // forward: w[co][ci][kh][kw]; backward: w[ci][co][2-kh][2-kw]
          const int widx = BWD ? ((ci * CO + co) * 3 + (2 - kh)) * 3 + (2 - kw)
                               : ((co * CI + ci) * 3 + kh) * 3 + kw;
          acc[co] = __builtin_fmaf(w[widx], xv, acc[co]);
        }
      }
    }
  }
  if (h < H && wc < W) {
    float* yb = y + (size_t)b * CO * plane + (size_t)h * W + wc;
#pragma unroll
    for (int co = 0; co < CO; co++) {
      float v = acc[co];
      if (relu && v < 0.0f) v = 0.0f;
      yb[(size_t)co * plane] = v;
    }
  }
}

// dW[co][ci][kh][kw] += sum_pixels g[b][co][p] * x[b][ci][p + (kh-1, kw-1)],
// g = dy * (mask > 0) when mask is given.  Block: one sample, a band of
// rows.  Thread t owns input channel ci = t mod CI and the PAIRS output
// channels co = (t / CI) * PAIRS + k (all nine taps each), so a pixel costs
// it 9 input reads and PAIRS gradient reads; it accumulates over the band's
// pixel tiles in registers and adds its partial sums once.  Each input
// channel's tile is padded to an odd length, so the lanes of a wave (32
// different channels at one pixel) read 32 different LDS banks.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv_dw(const float* __restrict__ x,
                                                const float* __restrict__ dy,
                                                const float* __restrict__ mask, int H, int W,
                                                int rows_per_block, float* __restrict__ dw) {
  constexpr int NP = CO * CI;                 // (co, ci) pairs
  constexpr int PAIRS = (NP + 255) / 256;
  constexpr int RH = 8, RW = 16;
  constexpr int XR = RW + 2, XS = (RH + 2) * XR + 1;   // odd channel stride
  static_assert(NP <= 256 || (NP % 256 == 0 && 256 % CI == 0), "thread owns one ci, PAIRS co");
  __shared__ float xs[CI][XS];
  __shared__ float gs[CO][RH][RW];
  const int b = blockIdx.y;
  const int hb = blockIdx.x * rows_per_block;
  const size_t plane = (size_t)H * W;
  const float* xb = x + (size_t)b * CI * plane;
  const float* gb = dy + (size_t)b * CO * plane;
  const float* mb = mask ? mask + (size_t)b * CO * plane : nullptr;
  const int ci = threadIdx.x % CI;
  const int co0 = (threadIdx.x / CI) * PAIRS;
  const bool owner = NP >= 256 || (int)threadIdx.x < NP;
  float acc[PAIRS][9];
#pragma unroll
  for (int k = 0; k < PAIRS; k++)
#pragma unroll
    for (int t = 0; t < 9; t++) acc[k][t] = 0.0f;
  for (int h0 = hb; h0 < hb + rows_per_block && h0 < H; h0 += RH) {
    for (int w0 = 0; w0 < W; w0 += RW) {
      __syncthreads();
      for (int e = threadIdx.x; e < CI * (RH + 2) * XR; e += 256) {
        const int c = e / ((RH + 2) * XR);
        const int rem = e - c * (RH + 2) * XR;
        const int yy = rem / XR, xx = rem - yy * XR;
        const int hh = h0 + yy - 1, ww = w0 + xx - 1;
        xs[c][rem] = (hh >= 0 && hh < H && ww >= 0 && ww < W)
                         ? xb[(size_t)c * plane + (size_t)hh * W + ww] : 0.0f;
      }
      for (int e = threadIdx.x; e < CO * RH * RW; e += 256) {
        const int c = e / (RH * RW);
        const int rem = e - c * RH * RW;
        const int yy = rem / RW, xx = rem - yy * RW;
        const int hh = h0 + yy, ww = w0 + xx;
        float v = 0.0f;
        if (hh < H && ww < W) {
          const size_t idx = (size_t)c * plane + (size_t)hh * W + ww;
          v = gb[idx];
          if (mb && !(mb[idx] > 0.0f)) v = 0.0f;
        }
        gs[c][yy][xx] = v;
      }
      __syncthreads();
      if (owner) {
#pragma unroll 1
        for (int yy = 0; yy < RH; yy++) {
#pragma unroll 2
          for (int xx = 0; xx < RW; xx++) {
            float xv[9];
#pragma unroll
            for (int kh = 0; kh < 3; kh++)
#pragma unroll
              for (int kw = 0; kw < 3; kw++) xv[kh * 3 + kw] = xs[ci][(yy + kh) * XR + xx + kw];
#pragma unroll
            for (int k = 0; k < PAIRS; k++) {
              const float g = gs[co0 + k][yy][xx];
#pragma unroll
              for (int t = 0; t < 9; t++) acc[k][t] = __builtin_fmaf(g, xv[t], acc[k][t]);
            }
          }
        }
      }
    }
  }
  if (owner) {
#pragma unroll
    for (int k = 0; k < PAIRS; k++)
#pragma unroll
      for (int t = 0; t < 9; t++)
        atomicAdd(&dw[((size_t)(co0 + k) * CI + ci) * 9 + t], acc[k][t]);
  }
}

// dB[co] = sum over batch and pixels of dy * (mask > 0), in two passes
// with a fixed summation order: part[co][g] over the g-th slice of the
// (batch, pixel) range (grid C x G), then db[co] = sum over g.
constexpr int kBiasSlices = 64;

__global__ __launch_bounds__(256) void k_bias_grad_part(const float* __restrict__ dy,
                                                        const float* __restrict__ mask, int B,
                                                        int C, int HW,
                                                        float* __restrict__ part) {
  __shared__ float red[4];
  const int co = blockIdx.x, g = blockIdx.y;
  const long long n = (long long)B * HW;
  const long long per = (n + kBiasSlices - 1) / kBiasSlices;
  const long long e0 = (long long)g * per, e1 = e0 + per < n ? e0 + per : n;
  float s = 0.0f;
  for (long long e = e0 + threadIdx.x; e < e1; e += 256) {
    const long long b = e / HW, p = e - b * HW;
    const size_t idx = ((size_t)b * C + co) * HW + (size_t)p;
    const float v = dy[idx];
    if (!mask || mask[idx] > 0.0f) s += v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)co * kBiasSlices + g] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_bias_grad_final(const float* __restrict__ part, int C, float* __restrict__ db) {
  const int co = blockIdx.x * blockDim.x + threadIdx.x;
  if (co >= C) return;
  float s = 0.0f;
  for (int g = 0; g < kBiasSlices; g++) s += part[(size_t)co * kBiasSlices + g];
  db[co] = s;
}

// mean BCE with logits: loss += sum(max(z,0) - z t + log1p(exp(-|z|))) / n,
// grad = (sigmoid(z) - t) / n
__global__ void k_bce(const float* __restrict__ z, const float* __restrict__ t, int64_t n,
                      float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float zi = z[i], ti = t[i];
    l = fmaxf(zi, 0.0f) - zi * ti + log1pf(expf(-fabsf(zi)));
    const float s = 1.0f / (1.0f + expf(-zi));
    grad[i] = (s - ti) / (float)n;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pbf, float* __restrict__ pw,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  const uint16_t q = f32_to_bf16(pi);
  pbf[i] = q;
  pw[i] = bf16_to_f32(q);
}

// bf16 copy of the master weights, and the same values widened to fp32
__global__ void k_to_bf16(const float* __restrict__ p, int n, uint16_t* __restrict__ out,
                          float* __restrict__ pw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint16_t q = f32_to_bf16(p[i]);
  out[i] = q;
  pw[i] = bf16_to_f32(q);
}

// ---------------------------------------------------------------------------
template <int CI, int CO, bool BWD>
static hipError_t conv_t(const float* x, const float* mask, const float* w, const float* b,
                         int B, int H, int W, int relu, float* y, hipStream_t st) {
  dim3 grid((W + TW - 1) / TW, (H + TH - 1) / TH, B);
  hipLaunchKernelGGL((k_conv<CI, CO, BWD>), grid, dim3(256), 0, st, x, mask, w, b, H, W, relu, y);
  return hipGetLastError();
}

hipError_t cnn_conv_forward(int ci, int co, const float* x, const float* w, const float* b,
                            int B, int H, int W, int relu, float* y, hipStream_t st) {
  if (ci == 1 && co == 32) return conv_t<1, 32, false>(x, nullptr, w, b, B, H, W, relu, y, st);
  if (ci == 32 && co == 32) return conv_t<32, 32, false>(x, nullptr, w, b, B, H, W, relu, y, st);
  if (ci == 32 && co == 1) return conv_t<32, 1, false>(x, nullptr, w, b, B, H, W, relu, y, st);
  return hipErrorInvalidValue;
}

// dx (ci_fwd channels) from dy (co_fwd channels); mask: the forward
// output of this layer when it had a ReLU (else nullptr)
hipError_t cnn_conv_backward_data(int ci, int co, const float* dy, const float* mask,
                                  const float* w, int B, int H, int W, float* dx,
                                  hipStream_t st) {
  if (ci == 32 && co == 32) return conv_t<32, 32, true>(dy, mask, w, nullptr, B, H, W, 0, dx, st);
  if (ci == 32 && co == 1) return conv_t<1, 32, true>(dy, mask, w, nullptr, B, H, W, 0, dx, st);
  if (ci == 1 && co == 32) return conv_t<32, 1, true>(dy, mask, w, nullptr, B, H, W, 0, dx, st);
  return hipErrorInvalidValue;
}

hipError_t cnn_conv_backward_weight(int ci, int co, const float* x, const float* dy,
                                    const float* mask, int B, int H, int W, float* dw,
                                    hipStream_t st) {
  const int rows = 16;
  dim3 grid((H + rows - 1) / rows, B);
  if (ci == 1 && co == 32)
    hipLaunchKernelGGL((k_conv_dw<1, 32>), grid, dim3(256), 0, st, x, dy, mask, H, W, rows, dw);
  else if (ci == 32 && co == 32)
    hipLaunchKernelGGL((k_conv_dw<32, 32>), grid, dim3(256), 0, st, x, dy, mask, H, W, rows, dw);
  else if (ci == 32 && co == 1)
    hipLaunchKernelGGL((k_conv_dw<32, 1>), grid, dim3(256), 0, st, x, dy, mask, H, W, rows, dw);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t cnn_bias_grad(const float* dy, const float* mask, int B, int C, int HW, float* part,
                         float* db, hipStream_t st) {
  hipLaunchKernelGGL(k_bias_grad_part, dim3(C, kBiasSlices), dim3(256), 0, st, dy, mask, B, C, HW,
                     part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bias_grad_final, dim3((C + 63) / 64), dim3(64), 0, st, part, C, db);
  return hipGetLastError();
}

int cnn_bias_part_floats(int C) { return C * kBiasSlices; }

hipError_t cnn_bce(const float* z, const float* t, int64_t n, float* grad, float* loss,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_bce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, z, t, n, grad,
                     loss);
  return hipGetLastError();
}

hipError_t cnn_adam(float* p, const float* g, float* m, float* v, uint16_t* pbf, float* pw, int n,
                    float lr, float b1, float b2, float eps, int step, hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p, g, m, v, pbf, pw, n, lr,
                     b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t cnn_to_bf16(const float* p, int n, uint16_t* out, float* pw, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, st, p, n, out, pw);
  return hipGetLastError();
}

}  // namespace fjk
