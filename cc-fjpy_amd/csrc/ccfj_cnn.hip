// ccfj_cnn.hip -- host side of the band-proposal CNN: parameters, forward,
// one training step (forward, backward, Adam), gradient export.
//
// Parameter vector (fp32 master, flat): w1[32][1][3][3] b1[32]
// w2[32][32][3][3] b2[32] w3[32][32][3][3] b3[32] w4[1][32][3][3] b4[1].
// The kernels read bf16 copies of the weights (biases stay fp32).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_cnn.h"
#include "fj_kernels.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

constexpr int kC = 32;
constexpr int kLayers = 4;
constexpr int kCin[kLayers] = {1, kC, kC, kC};
constexpr int kCout[kLayers] = {kC, kC, kC, 1};

struct Offsets {
  size_t w[kLayers], b[kLayers], total;
  Offsets() {
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) {
      w[l] = o;
      o += (size_t)kCout[l] * kCin[l] * 9;
      b[l] = o;
      o += kCout[l];
    }
    total = o;
  }
};
const Offsets kOff;

}  // namespace

struct ccfj_cnn {
  ccfj_ctx* ctx = nullptr;
  int device = 0;
  int H = 0, W = 0, max_batch = 0;
  int step = 0;
  std::vector<void*> owned;
  float* p = nullptr;       // fp32 master
  float* g = nullptr;       // gradients
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  uint16_t* pb = nullptr;   // bf16 copy (weights and biases)
  float* pw = nullptr;      // the bf16 values widened to fp32 (the convs read the weights)
  float* part = nullptr;    // bias-gradient partial sums
  float* act[kLayers] = {};  // layer outputs [B][cout][H][W]
  float* gA = nullptr;      // gradient buffers [B][32][H][W]
  float* gB = nullptr;
  float* loss = nullptr;
};

namespace {

int alloc(ccfj_cnn* n, size_t bytes, void** out) {
  int rc = ccfj_internal::device_alloc(out, bytes);
  if (!rc) n->owned.push_back(*out);
  return rc;
}

int forward(ccfj_cnn* n, const float* x, int B, hipStream_t st) {
  const float* in = x;
  for (int l = 0; l < kLayers; l++) {
    hipError_t e = fjk::cnn_conv_forward(kCin[l], kCout[l], in, n->pw + kOff.w[l], n->p + kOff.b[l],
                                         B, n->H, n->W, l < kLayers - 1 ? 1 : 0, n->act[l], st);
    if (e != hipSuccess) return hip_fail(e, "conv forward");
    in = n->act[l];
  }
  return CCFJ_OK;
}

// gradients of the mean BCE w.r.t. all parameters into n->g; loss into n->loss
int backward(ccfj_cnn* n, const float* x, const float* target, int B, hipStream_t st) {
  const int HW = n->H * n->W;
  hipError_t e = fjk::launch_fill(n->g, kOff.total, 0.0f, st);
  if (e == hipSuccess) e = fjk::launch_fill(n->loss, 1, 0.0f, st);
  if (e != hipSuccess) return hip_fail(e, "fill kernel launch");
  // dL/dlogits -> gA (1 channel)
  e = fjk::cnn_bce(n->act[kLayers - 1], target, (int64_t)B * HW, n->gA, n->loss, st);
  if (e != hipSuccess) return hip_fail(e, "bce");
  float* gcur = n->gA;   // gradient w.r.t. the output of layer l (before its ReLU mask)
  float* gnext = n->gB;
  for (int l = kLayers - 1; l >= 0; l--) {
    const float* mask = l < kLayers - 1 ? n->act[l] : nullptr;  // ReLU layers
    const float* in = l > 0 ? n->act[l - 1] : x;
    e = fjk::cnn_conv_backward_weight(kCin[l], kCout[l], in, gcur, mask, B, n->H, n->W,
                                      n->g + kOff.w[l], st);
    if (e != hipSuccess) return hip_fail(e, "conv backward weight");
    e = fjk::cnn_bias_grad(gcur, mask, B, kCout[l], HW, n->part, n->g + kOff.b[l], st);
    if (e != hipSuccess) return hip_fail(e, "bias grad");
    if (l > 0) {
      e = fjk::cnn_conv_backward_data(kCin[l], kCout[l], gcur, mask, n->pw + kOff.w[l], B, n->H,
                                      n->W, gnext, st);
      if (e != hipSuccess) return hip_fail(e, "conv backward data");
      float* t = gcur;
      gcur = gnext;
      gnext = t;
    }
  }
  return CCFJ_OK;
}

int check_batch(ccfj_cnn* n, int B) {
  if (!n) return fail(CCFJ_EINVAL, "null network");
  if (B < 1 || B > n->max_batch) return fail(CCFJ_EINVAL, "batch %d outside 1..%d", B, n->max_batch);
  hipError_t e = hipSetDevice(n->device);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipSetDevice");
}

}  // namespace

extern "C" {

int ccfj_cnn_param_count(void) { return (int)kOff.total; }

int ccfj_cnn_create(ccfj_ctx* ctx, int H, int W, int max_batch, unsigned long long seed,
                    ccfj_cnn** out) {
  if (!ctx || !out) return fail(CCFJ_EINVAL, "null argument");
  *out = nullptr;
  if (H < 1 || W < 1 || max_batch < 1) return fail(CCFJ_EINVAL, "bad network shape");
  ccfj_cnn* n = new ccfj_cnn();
  n->ctx = ctx;
  n->device = ccfj_internal::ctx_device(ctx);
  n->H = H;
  n->W = W;
  n->max_batch = max_batch;
  hipError_t e = hipSetDevice(n->device);
  if (e != hipSuccess) {
    delete n;
    return hip_fail(e, "hipSetDevice");
  }
  const size_t P = kOff.total;
  const size_t A = (size_t)max_batch * kC * H * W;
  void* q;
  int rc = CCFJ_OK;
  if (!rc && !(rc = alloc(n, sizeof(float) * P, &q))) n->p = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * P, &q))) n->g = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * P, &q))) n->m = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * P, &q))) n->v = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(uint16_t) * P, &q))) n->pb = (uint16_t*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * P, &q))) n->pw = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * fjk::cnn_bias_part_floats(kC), &q)))
    n->part = (float*)q;
  for (int l = 0; l < kLayers && !rc; l++)
    if (!(rc = alloc(n, sizeof(float) * (size_t)max_batch * kCout[l] * H * W, &q))) n->act[l] = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * A, &q))) n->gA = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float) * A, &q))) n->gB = (float*)q;
  if (!rc && !(rc = alloc(n, sizeof(float), &q))) n->loss = (float*)q;
  if (rc) {
    ccfj_cnn_destroy(n);
    return rc;
  }
  // He-normal initialisation, zero biases
  std::vector<float> h(P, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; l++) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * kCin[l])));
    for (size_t k = 0; k < (size_t)kCout[l] * kCin[l] * 9; k++) h[kOff.w[l] + k] = nd(rng);
  }
  rc = ccfj_cnn_set_params(n, h.data());
  if (rc) {
    ccfj_cnn_destroy(n);
    return rc;
  }
  *out = n;
  return CCFJ_OK;
}

int ccfj_cnn_destroy(ccfj_cnn* n) {
  if (!n) return CCFJ_OK;
  (void)hipSetDevice(n->device);
  (void)hipDeviceSynchronize();
  for (void* q : n->owned) (void)hipFree(q);
  delete n;
  return CCFJ_OK;
}

int ccfj_cnn_set_params(ccfj_cnn* n, const float* params) {
  if (!n || !params) return fail(CCFJ_EINVAL, "null argument");
  hipError_t e = hipSetDevice(n->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(n->ctx);
  const size_t P = kOff.total;
  e = hipMemcpyAsync(n->p, params, sizeof(float) * P, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = fjk::launch_fill(n->m, P, 0.0f, st);
  if (e == hipSuccess) e = fjk::launch_fill(n->v, P, 0.0f, st);
  if (e == hipSuccess) e = fjk::cnn_to_bf16(n->p, (int)P, n->pb, n->pw, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  n->step = 0;
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "set params");
}

int ccfj_cnn_get_params(ccfj_cnn* n, float* params) {
  if (!n || !params) return fail(CCFJ_EINVAL, "null argument");
  hipError_t e = hipSetDevice(n->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ccfj_internal::ctx_stream(n->ctx));
  if (e == hipSuccess) e = hipMemcpy(params, n->p, sizeof(float) * kOff.total, hipMemcpyDeviceToHost);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "get params");
}

int ccfj_cnn_forward(ccfj_cnn* n, const float* x, int B, float* logits, void* stream) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ccfj_internal::ctx_stream(n->ctx);
  if ((rc = forward(n, x, B, st))) return rc;
  hipError_t e = hipMemcpyAsync(logits, n->act[kLayers - 1], sizeof(float) * (size_t)B * n->H * n->W,
                                hipMemcpyDeviceToDevice, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipMemcpyAsync");
}

int ccfj_cnn_gradients(ccfj_cnn* n, const float* x, const float* target, int B, float* grads,
                       float* loss) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  hipStream_t st = ccfj_internal::ctx_stream(n->ctx);
  if ((rc = forward(n, x, B, st)) || (rc = backward(n, x, target, B, st))) return rc;
  hipError_t e = hipMemcpyAsync(grads, n->g, sizeof(float) * kOff.total, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "gradients");
}

int ccfj_cnn_train_step(ccfj_cnn* n, const float* x, const float* target, int B, float lr,
                        float* loss) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  hipStream_t st = ccfj_internal::ctx_stream(n->ctx);
  if ((rc = forward(n, x, B, st)) || (rc = backward(n, x, target, B, st))) return rc;
  n->step++;
  hipError_t e = fjk::cnn_adam(n->p, n->g, n->m, n->v, n->pb, n->pw, (int)kOff.total, lr, 0.9f, 0.999f,
                               1e-8f, n->step, st);
  if (e != hipSuccess) return hip_fail(e, "adam");
  if (loss) {
    e = hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "loss copy");
  }
  return CCFJ_OK;
}

}  // extern "C"
