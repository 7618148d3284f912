// kernels.h -- launchers of the HIP kernels (internal to libhipdvr.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "march_common.h"

namespace hipdvr {

// Ray-generation constants of the unit-cube proxy, texture space.
struct RayGen {
    float o[3];     // camera position
    float f[3];     // view direction
    float r[3];     // right * tan(fovy/2) * aspect
    float u[3];     // up * tan(fovy/2)
    float near_t;   // near-plane ray parameter
    int32_t w, h;
    int32_t quantize;
};

void launch_march_ref(const MarchArgs& a, int dtype, int filter, int classif, bool phong,
                      bool count, hipStream_t st);
// compositing modes other than DVR (TF classification), reference layout
void launch_march_modes(const MarchArgs& a, int dtype, int filter, bool phong, hipStream_t st);
void launch_march_fast(const MarchArgs& a, int dtype, int filter, bool phong, bool count, bool wide,
                       hipStream_t st);
// shade: 0 none, 1 Phong (interior gradient, approximate lighting), 2 Phong
// bit-identical to the reference kernel; batch: samples fetched together per
// ray (2 or 4)
void launch_march_cell(const MarchArgs& a, int dtype, int shade, bool count, bool wide, int batch,
                       hipStream_t st);
// The slab-staged march (march_slab.h): linear filter, TF classification,
// DVR; shade as for the cell kernel.  Reads the linear volume.
void launch_march_slab(const MarchArgs& a, int dtype, int shade, bool count, hipStream_t st);
// Samples in flight per ray by default, measured on 512^3 / 1024^2: four
// unshaded (1.00 ms against 1.04 with two and 1.28 with one); two shaded
// (1.88 ms against 2.01 with four and 2.18 with one, Phong within
// tolerance; the exact-shading path's six full lookups hold more registers
// too).
constexpr int default_batch(int shade) { return shade ? 2 : 4; }

// Unshaded marches run five workgroups per CU: their dynamic LDS is padded so
// that a workgroup holds 32 KiB of the CU's 160 KiB.  With the seven their
// VGPRs allow, the rays in flight evict each other's records from the caches
// (512^3 / 1024^2: plain 1.07 against 1.00 ms; learned 1974 against 2053
// frames/s).  Returns the dynamic LDS bytes for a kernel whose static LDS is
// the TF table and the opacity table and that needs `dyn` bytes itself.
constexpr size_t kUnshadedLdsBytes = 32768;
constexpr size_t kMarchStaticLds = sizeof(uint32_t) * kMaxTfWidth + sizeof(float) * 256;
constexpr size_t unshaded_dyn_lds(size_t dyn) {
    return dyn < kUnshadedLdsBytes - kMarchStaticLds ? kUnshadedLdsBytes - kMarchStaticLds : dyn;
}

// Zero n 64-bit counters with a kernel (graph-capture safe, unlike a memset node).
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
// The plain re-march of a sparse pixel grid (a.grid_w x a.grid_h, a.stride,
// a.pix_offset) in ray segments (sparse.hip); seg: sparse_segment_bytes().
size_t sparse_segment_bytes(uint32_t rays);
void launch_sparse_remarch(const MarchArgs& a, int dtype, int shade, bool wide, float4* seg, hipStream_t st);
// Streaming copy of n 16-byte vectors (bandwidth calibration).
void launch_copy16(void* dst, const void* src, size_t n, hipStream_t st);
void launch_entry_exit(const RayGen& g, float4* entry, float4* exit, hipStream_t st);

}  // namespace hipdvr
