// mlp.h -- the stepping network (internal): sizes, device parameter views,
// launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hipdvr {

constexpr int kMlpIn = 16;       // brick/TF/octant features (zero padded)
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;       // skip logit, log2 step multiplier
constexpr int kMlpTargets = 3;   // skip label, log2 multiplier label, multiplier weight
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kMlpOut + kMlpOut;

// BF16 weights and FP32 biases the kernels read.
struct MlpDevParams {
    short* w1;   // [16][64]
    short* w2;   // [64][64]
    short* w3;   // [64][16], columns >= 2 zero
    float* b;    // b1[64] b2[64] b3[2]
};

void launch_mlp_pack(const float* master, const MlpDevParams& p, hipStream_t st);
void launch_mlp_forward(const MlpDevParams& p, const float* x, uint32_t rows, float* y, hipStream_t st);
void launch_mlp_grad(const MlpDevParams& p, const float* x, const float* tgt, uint32_t rows, float* grad,
                     float* loss, hipStream_t st);
void launch_mlp_adam(float* w, float* m, float* v, const float* g, float lr, int step, hipStream_t st);

}  // namespace hipdvr
