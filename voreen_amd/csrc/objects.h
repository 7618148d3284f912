// objects.h -- transfer-function object and the render helpers shared by
// the plain and learned paths (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"
#include "march_common.h"
#include "volume.h"

struct hipdvr_tf {
    hipdvr_ctx* ctx = nullptr;
    uint64_t serial = 0;             // unique per TF object (occupancy cache key)
    uint32_t width = 0;
    float domain_lo = 0.f, domain_hi = 1.f;
    uint32_t* d_rgba = nullptr;
    uint32_t* d_alpha_prefix = nullptr;   // [width+1] count of non-zero alpha bytes below i
    std::vector<uint8_t> host_rgba;
    float4* d_preint = nullptr;
    uint32_t preint_n = 0;
    // 2D TF (hipdvr_tf2d_create): float RGBA table, its domains, and the
    // opacity-corrected copy on the device for the step it was made for
    bool is2d = false;
    uint32_t w2d = 0, h2d = 0;
    float dom2d[4] = {0.f, 1.f, 0.f, 1.f};
    std::vector<float> host2d;
    std::vector<float> host2d_corr;
    float4* d_tf2d = nullptr;
    float tf2d_step = -1.f;
};

namespace hipdvr {

void fill_alpha_table(float step, float* out);
void validate_params(const hipdvr_params* p, const hipdvr_tf* tf);
void fill_march_args(MarchArgs& a, const hipdvr_volume* vol, const hipdvr_tf* tf, const hipdvr_params* p,
                     const float* d_entry, const float* d_exit, float* d_out, uint32_t w, uint32_t h,
                     unsigned long long* counters);
void ensure_occupancy(hipdvr_ctx* ctx, hipdvr_volume* vol, const hipdvr_tf* tf);
// Opacity-corrected 2D table for this step on the device (host powf, as
// cpuraycaster.cpp:289 evaluates it per sample).
void ensure_tf2d(hipdvr_tf* tf, float step, hipStream_t st);
// TF colour per pre-integration table index (RGBA float x n); false on bad keys.
bool preint_index_colors(const hipdvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo, float thr_hi,
                         int32_t alpha_mode, uint32_t n, float* tfb);

}  // namespace hipdvr
