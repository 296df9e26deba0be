// fj_cnn.h -- kernels of the band-proposal CNN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fjk {

hipError_t cnn_conv_forward(int ci, int co, const float* x, const uint16_t* w, const float* b,
                            int B, int H, int W, int relu, float* y, hipStream_t st);
hipError_t cnn_conv_backward_data(int ci, int co, const float* dy, const float* mask,
                                  const uint16_t* w, int B, int H, int W, float* dx,
                                  hipStream_t st);
hipError_t cnn_conv_backward_weight(int ci, int co, const float* x, const float* dy,
                                    const float* mask, int B, int H, int W, float* dw,
                                    hipStream_t st);
hipError_t cnn_bias_grad(const float* dy, const float* mask, int B, int C, int HW, float* db,
                         hipStream_t st);
hipError_t cnn_bce(const float* z, const float* t, int64_t n, float* grad, float* loss,
                   hipStream_t st);
hipError_t cnn_adam(float* p, const float* g, float* m, float* v, uint16_t* pbf, int n, float lr,
                    float b1, float b2, float eps, int step, hipStream_t st);
hipError_t cnn_to_bf16(const float* p, int n, uint16_t* out, hipStream_t st);

}  // namespace fjk
