// preint.hip -- pre-integration tables built on the device.
//
// PreIntegrationTable::computeTable (src/core/datastructures/transfunc/1d/
// preintegrationtable.cpp:67-204), both variants, one lane per (sb, sf)
// entry: the exact variant composites the TF from sf to sb with the opacity
// correction exponent step*200/(|sb-sf|+1) and early termination at 0.95,
// then un-premultiplies; the integral variant differences prefix sums
// (alpha = 1 - exp(-dA*step*200/ds)).  The TF colours per table index and the
// integral prefix sums are made on the host in the reference's float order
// (hipdvr_preint_build, tf.cpp); the device evaluates pow/exp with its own
// library functions (within an ulp of the host's).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/hipdvr.h"
#include "ctx.h"
#include "errors.h"
#include "objects.h"

namespace hipdvr {

namespace {

__device__ __forceinline__ float dmaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float dminf(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float clamp01(float v) { return dminf(dmaxf(v, 0.f), 1.f); }

__global__ __launch_bounds__(256) void preint_exact_kernel(const float4* __restrict__ tfb, int n, float step,
                                                           float4* __restrict__ table) {
    const int sf = blockIdx.x * blockDim.x + threadIdx.x;
    const int sb = blockIdx.y;
    if (sf >= n) return;
    float4 out;
    if (sb != sf) {
        const float scale = 1.0f / (fabsf((float)(sb - sf)) + 1);
        const int incr = (sb < sf) ? -1 : 1;
        float r = 0.f, g = 0.f, b = 0.f, a = 0.f;
        // the reference compares the float alpha with the double 0.95
        // (preintegrationtable.cpp:115); 0.95f < 0.95 < the next float up, so
        // that test is a <= 0.95f, without double arithmetic
        for (int s = sf; (incr == 1 ? s <= sb : s >= sb) && (a <= 0.95f); s += incr) {
            const float4 cc = tfb[s];
            float ca = cc.w;
            if (ca > 0.0f) {
                ca = 1.f - powf(1.f - ca, step * 200.0f * scale);
                const float w = (1.0f - a) * ca;
                const float nr = r + w * cc.x;
                const float ng = g + w * cc.y;
                const float nb = b + w * cc.z;
                a = a + (1.0f - a) * ca;
                r = nr;
                g = ng;
                b = nb;
            }
        }
        const float den = dmaxf(a, 0.001f);
        out = make_float4(clamp01(r / den), clamp01(g / den), clamp01(b / den), clamp01(a));
    } else {
        const float4 cc = tfb[sf];
        out = make_float4(clamp01(cc.x), clamp01(cc.y), clamp01(cc.z), 1.f - powf(1.f - clamp01(cc.w), step * 200.0f));
    }
    table[(size_t)sb * n + sf] = out;
}

__global__ __launch_bounds__(256) void preint_integral_kernel(const float4* __restrict__ tfb,
                                                              const float4* __restrict__ intf, int n, float step,
                                                              float4* __restrict__ table) {
    const int sf = blockIdx.x * blockDim.x + threadIdx.x;
    const int sb = blockIdx.y;
    if (sf >= n) return;
    const int smin = sb < sf ? sb : sf;
    const int smax = sb < sf ? sf : sb;
    float col[4];
    if (smax != smin) {
        const float factor = step * 200.f / (float)(smax - smin);
        const float4 hi = intf[smax], lo = intf[smin];
        col[0] = (hi.x - lo.x) * factor;
        col[1] = (hi.y - lo.y) * factor;
        col[2] = (hi.z - lo.z) * factor;
        col[3] = 1.f - expf(-(hi.w - lo.w) * factor);
        const float den = dmaxf(col[3], 0.001f);
        for (int c = 0; c < 3; ++c) col[c] /= den;
    } else {
        const float4 cc = tfb[smin];
        col[0] = cc.x;
        col[1] = cc.y;
        col[2] = cc.z;
        col[3] = 1.f - powf(1.f - cc.w, step * 200.0f);
    }
    table[(size_t)sb * n + sf] = make_float4(clamp01(col[0]), clamp01(col[1]), clamp01(col[2]), clamp01(col[3]));
}

}  // namespace

}  // namespace hipdvr

using namespace hipdvr;

extern "C" int hipdvr_tf_preint_device(hipdvr_tf* tf, const hipdvr_tf_key* keys, int32_t nkeys, float gamma,
                                       float thr_lo, float thr_hi, int32_t alpha_mode, uint32_t n, float step,
                                       int32_t use_integral) {
    if (!tf) return HIPDVR_ERR_INVALID;
    return ctx_guarded(tf->ctx, [&] {
        if (n < 2 || n > 4096 || !(step > 0.f) || nkeys < 0 || (nkeys > 0 && !keys))
            fail(HIPDVR_ERR_INVALID, "bad pre-integration arguments");
        // TF colours per table index: the host builder's first stage
        std::vector<float> tfb((size_t)n * 4);
        if (!preint_index_colors(keys, nkeys, gamma, thr_lo, thr_hi, alpha_mode, n, tfb.data()))
            fail(HIPDVR_ERR_INVALID, "invalid transfer-function keys");
        std::vector<float> intf;
        if (use_integral) {
            intf.assign((size_t)n * 4, 0.f);
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t i = 0; i < n; ++i) {
                const float* cc = &tfb[4 * (size_t)i];
                if (cc[3] > 0.0f) {
                    acc[0] += cc[0] * cc[3];
                    acc[1] += cc[1] * cc[3];
                    acc[2] += cc[2] * cc[3];
                    acc[3] += cc[3];
                }
                for (int c = 0; c < 4; ++c) intf[4 * (size_t)i + c] = acc[c];
            }
        }
        ctx_set_device(tf->ctx);
        hipStream_t st = ctx_stream(tf->ctx);
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        float4* d_tab = nullptr;
        float4* d_tfb = nullptr;
        float4* d_int = nullptr;
        auto cleanup = [&] {
            if (d_tfb) (void)hipFree(d_tfb);
            if (d_int) (void)hipFree(d_int);
        };
        try {
            hip_check(hipMalloc((void**)&d_tab, (size_t)n * n * sizeof(float4)), "hipMalloc(preint)");
            hip_check(hipMalloc((void**)&d_tfb, (size_t)n * sizeof(float4)), "hipMalloc(preint)");
            hip_check(hipMemcpyAsync(d_tfb, tfb.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, st),
                      "upload");
            const dim3 grid((n + 255) / 256, n);
            if (use_integral) {
                hip_check(hipMalloc((void**)&d_int, (size_t)n * sizeof(float4)), "hipMalloc(preint)");
                hip_check(hipMemcpyAsync(d_int, intf.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, st),
                          "upload");
                hipLaunchKernelGGL(preint_integral_kernel, grid, dim3(256), 0, st, d_tfb, d_int, (int)n, step, d_tab);
            } else {
                hipLaunchKernelGGL(preint_exact_kernel, grid, dim3(256), 0, st, d_tfb, (int)n, step, d_tab);
            }
            hip_check(hipGetLastError(), "pre-integration kernel launch");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            if (d_tab) (void)hipFree(d_tab);
            cleanup();
            throw;
        }
        cleanup();
        if (tf->d_preint) (void)hipFree(tf->d_preint);
        tf->d_preint = d_tab;
        tf->preint_n = n;
    });
}
