// entry_exit.hip -- entry/exit points of the unit-cube proxy, one lane per
// pixel.  The GPU counterpart of MeshEntryExitPoints for a bounding-box proxy
// (modules/base/processors/entryexitpoints/meshentryexitpoints.cpp:188-273):
// entry = first front-face hit in texture coordinates, exit = last back-face
// hit, both with w = 1; a camera inside the cube starts at the near plane
// (modules/base/glsl/eep_insidevolume.frag:44-76); missed pixels are exactly 0.
// Compiled with -ffp-contract=off so the float sequence is the one of the
// host-side restatement used to check it.
#include "kernels.h"

namespace hipdvr {

__device__ __forceinline__ float clamp01(float v) {
    v = (v < 0.f) ? 0.f : v;
    return (1.f < v) ? 1.f : v;
}

__global__ __launch_bounds__(256) void entry_exit_kernel(const RayGen g, float4* __restrict__ entry,
                                                         float4* __restrict__ exit) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.w || y >= g.h) return;
    const float nx = (2.0f * ((float)x + 0.5f)) / (float)g.w - 1.0f;
    const float ny = (2.0f * ((float)y + 0.5f)) / (float)g.h - 1.0f;
    float d[3];
    float tn = -INFINITY, tf = INFINITY;
    bool miss = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        d[i] = (g.f[i] + nx * g.r[i]) + ny * g.u[i];
        if (d[i] != 0.0f) {
            const float inv = 1.0f / d[i];
            const float ta = (0.0f - g.o[i]) * inv;
            const float tb = (1.0f - g.o[i]) * inv;
            tn = fmaxf(tn, fminf(ta, tb));
            tf = fminf(tf, fmaxf(ta, tb));
        } else if (g.o[i] < 0.0f || g.o[i] > 1.0f) {
            miss = true;
        }
    }
    const float ts = fmaxf(tn, g.near_t);
    const size_t p = (size_t)y * g.w + x;
    if (miss || !(tf > ts)) {
        entry[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        exit[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    float en[3], ex[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float a = clamp01(g.o[i] + ts * d[i]);
        float b = clamp01(g.o[i] + tf * d[i]);
        if (g.quantize) {
            a = floorf(a * 65535.0f + 0.5f) / 65535.0f;
            b = floorf(b * 65535.0f + 0.5f) / 65535.0f;
        }
        en[i] = a;
        ex[i] = b;
    }
    entry[p] = make_float4(en[0], en[1], en[2], 1.0f);
    exit[p] = make_float4(ex[0], ex[1], ex[2], 1.0f);
}

void launch_entry_exit(const RayGen& g, float4* entry, float4* exit, hipStream_t st) {
    dim3 grid((g.w + 15) / 16, (g.h + 15) / 16);
    hipLaunchKernelGGL(entry_exit_kernel, grid, dim3(256), 0, st, g, entry, exit);
}

}  // namespace hipdvr
