// learned.h -- learned stepping internals: brick size, kernel arguments,
// launchers (see learned_march.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipdvr_learned.h"
#include "march_common.h"

namespace hipdvr {

constexpr uint32_t kLBrickLog2 = 4;
constexpr uint32_t kLBrick = 1u << kLBrickLog2;   // 16^3 bricks
constexpr int kLFeatures = 16;
constexpr int kLMaxMult = 8;                      // largest step multiplier

struct FeatureArgs {
    const float4* stats;        // per brick: min, max, mean, var (normalized)
    uint32_t rows;              // bricks * 8
    float rwm_scale, rwm_offset, dom_lo, dom_hi;
    int32_t tf_width;
    const float* alpha;         // [w] opacity-corrected alpha per TF index (base step)
    const float* alpha_prefix;  // [w+1] prefix sums of alpha
    const uint32_t* nz_prefix;  // [w+1] prefix counts of non-zero alpha bytes
    float* x;                   // [rows][16] features out
    uint8_t* occ16;             // [bricks] exact occupancy at 16^3 (written for octant 0)
};

struct LearnedArgs {
    MarchArgs m;                // volume, TF, images, Phong, step (base)
    int32_t u16;
    uint32_t nb16[3];
    const uint8_t* table;       // [bricks*8] bit 7 skip, bits 0..1 log2 multiplier
    const uint8_t* occ16;       // [bricks] exact occupancy
    const float* alpha_m;       // [8][256] opacity correction for step * 1 .. 8 (device)
};

void launch_brick_stats(const void* linear, int dtype, const uint32_t dims[3], const uint32_t nb[3],
                        float4* out, hipStream_t st);
void launch_features(const FeatureArgs& a, hipStream_t st);
void launch_table(const float* y, uint32_t rows, uint8_t* table, float skip_logit, int max_log2,
                  hipStream_t st);
void launch_fill_table(uint8_t* table, uint32_t rows, uint8_t v, hipStream_t st);
void launch_merge_occ(uint8_t* table, const uint8_t* occ16, uint32_t rows, hipStream_t st);
// count[0] = rows where the table departs from the plain march
void launch_table_gain(const uint8_t* table, const uint8_t* occ16, uint32_t rows, unsigned int* count,
                       hipStream_t st);
void launch_trace(const LearnedArgs& la, float* err, unsigned int* visits, hipStream_t st);
void launch_targets(const float* err, const unsigned int* visits, const float* feat, uint32_t rows,
                    float tau_skip, float tau_step, float* out_x, float* out_t, unsigned int* count,
                    uint32_t capacity, hipStream_t st);
void launch_learned(const LearnedArgs& la, bool phong, bool count, bool wide, hipStream_t st);

// For batched launches (batch.hip): the judge's sparse grid of a w x h image
// (host arithmetic), the options with their defaults, the stepping fields of
// a LearnedArgs for one job (table, occupancy, opacity tables for the job's
// step; la.m is left to the caller), the opacity tables remade for a step
// (nothing when they hold it), and the TF of the last prepare (NULL before).
struct JudgeGrid {
    uint32_t stride, block, offs;
    int w, h;
};
JudgeGrid learned_judge_grid(uint32_t w, uint32_t h, const hipdvr_learned_opts& opts);
hipdvr_learned_opts learned_options(const hipdvr_learned_opts* o);
void learned_stepping_args(hipdvr_learned* l, const hipdvr_params* p, LearnedArgs& la);
const hipdvr_volume* learned_volume(const hipdvr_learned* l);
void learned_ensure_step(hipdvr_learned* l, float step);
// whether the last prepared table departs from the plain march anywhere
// (false: the learned frame would be the plain frame at a higher cost)
bool learned_gains(const hipdvr_learned* l);
const hipdvr_tf* learned_prepared_tf(const hipdvr_learned* l);
void launch_judge(const float4* learned, int w, int h, const float4* sparse, int sw, int sh, int stride,
                  int block, int offs, float* sse, int* flag, float mse_max, hipStream_t st);
// The judges of n views in one launch (jobs in device memory); denom is the
// sparse grid's 4 * sw * sh, as launch_judge passes it.
struct JudgeJob {
    const float4* learned;
    const float4* sparse;
    float* sse;
    int* flag;
    int32_t w, h, sw, sh, stride, block, offs;
    float denom, mse_max;
};
void launch_judge_multi(const JudgeJob* jobs, uint32_t n, hipStream_t st);

}  // namespace hipdvr
