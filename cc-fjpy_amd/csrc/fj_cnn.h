// fj_cnn.h -- kernels of the band-proposal CNN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fjk {

// w: the bf16 weights widened to fp32
hipError_t cnn_conv_forward(int ci, int co, const float* x, const float* w, const float* b,
                            int B, int H, int W, int relu, float* y, hipStream_t st);
hipError_t cnn_conv_backward_data(int ci, int co, const float* dy, const float* mask,
                                  const float* w, int B, int H, int W, float* dx,
                                  hipStream_t st);
hipError_t cnn_conv_backward_weight(int ci, int co, const float* x, const float* dy,
                                    const float* mask, int B, int H, int W, float* dw,
                                    hipStream_t st);
// part: cnn_bias_part_floats(C) floats of workspace
hipError_t cnn_bias_grad(const float* dy, const float* mask, int B, int C, int HW, float* part,
                         float* db, hipStream_t st);
int cnn_bias_part_floats(int C);
hipError_t cnn_bce(const float* z, const float* t, int64_t n, float* grad, float* loss,
                   hipStream_t st);
// pbf: bf16 copy of p; pw: the same values widened to fp32 (what the convs read)
hipError_t cnn_adam(float* p, const float* g, float* m, float* v, uint16_t* pbf, float* pw, int n,
                    float lr, float b1, float b2, float eps, int step, hipStream_t st);
hipError_t cnn_to_bf16(const float* p, int n, uint16_t* out, float* pw, hipStream_t st);

}  // namespace fjk
