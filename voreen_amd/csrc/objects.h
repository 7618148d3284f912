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
};

namespace hipdvr {

void fill_alpha_table(float step, float* out);
void validate_params(const hipdvr_params* p, const hipdvr_tf* tf);
void fill_march_args(MarchArgs& a, const hipdvr_volume* vol, const hipdvr_tf* tf, const hipdvr_params* p,
                     const float* d_entry, const float* d_exit, float* d_out, uint32_t w, uint32_t h,
                     unsigned long long* counters);
void ensure_occupancy(hipdvr_ctx* ctx, hipdvr_volume* vol, const hipdvr_tf* tf);
// TF colour per pre-integration table index (RGBA float x n); false on bad keys.
bool preint_index_colors(const hipdvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo, float thr_hi,
                         int32_t alpha_mode, uint32_t n, float* tfb);

}  // namespace hipdvr
