// synth_device.hip -- seeded synthetic volumes generated on the device
// (include/hipdvr_synth.h).  The large-volume stress workload is 8 GiB of
// uint16: it is made where it is marched, with no host copy.
//
// Turbulence: sum over 6 octaves of 2^-o |value noise| at frequency
// 6 * 2^o lattice cells across the largest axis (normalized voxel centres),
// divided by the sum of the weights, quantized to [0, 65535].  Then the
// value below which `empty_fraction` of the voxels lie (a 4096-bin histogram
// of the top 12 bits, boundary nearest the target) becomes the threshold:
// voxels under it are set to 0, the rest rescaled from [threshold, top of
// the highest occupied bin] onto [1, 65535].
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/hipdvr.h"
#include "../../include/hipdvr_synth.h"
#include "ctx.h"
#include "errors.h"
#include "kernels.h"
#include "synth_scratch.h"

// Histogram and zero counter of hipdvr_synth_turbulence_device_noalloc:
// 4096 bins + 1 counter of uint64 (32,776 bytes), whatever the volume size.
// One copy per GPU, held until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device
// (g_scratch below), never through this name.
static __device__ __align__(256) unsigned long long scratch[4096 + 1];

namespace hipdvr {

namespace {

constexpr int kOctaves = 6;
constexpr int kHistBins = 4096;   // top 12 bits of a uint16 value

// 32-bit integer mixer (xor-shift / multiply rounds).
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Lattice value in [-1, 1) of integer point (x, y, z) for one octave seed.
__device__ __forceinline__ float lattice(uint32_t seed, int x, int y, int z) {
    const uint32_t h = mix32((uint32_t)x * 0x8da6b343u + mix32((uint32_t)y * 0xd8163841u +
                                                            mix32((uint32_t)z * 0xcb1ab31fu + seed)));
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

__device__ __forceinline__ float smooth(float t) { return t * t * (3.0f - 2.0f * t); }

__device__ float value_noise(uint32_t seed, float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const float tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        float row = 0.0f;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float a = lattice(seed, ix, iy + j, iz + k);
            const float b = lattice(seed, ix + 1, iy + j, iz + k);
            row += (j ? ty : 1.0f - ty) * (a + (b - a) * tx);
        }
        acc += (k ? tz : 1.0f - tz) * row;
    }
    return acc;
}

__global__ __launch_bounds__(256) void turbulence_kernel(uint16_t* __restrict__ out, uint32_t nx, uint32_t ny,
                                                         uint64_t n, float inv_m, uint32_t seed) {
    uint32_t oseed[kOctaves];
#pragma unroll
    for (int o = 0; o < kOctaves; ++o) oseed[o] = mix32(seed + 0x9e3779b9u * (uint32_t)(o + 1));
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t x = (uint32_t)(i % nx);
        const uint64_t r = i / nx;
        const uint32_t y = (uint32_t)(r % ny), z = (uint32_t)(r / ny);
        const float px = ((float)x + 0.5f) * inv_m, py = ((float)y + 0.5f) * inv_m, pz = ((float)z + 0.5f) * inv_m;
        float sum = 0.0f, w = 1.0f, freq = 6.0f, norm = 0.0f;
#pragma unroll
        for (int o = 0; o < kOctaves; ++o) {
            sum += w * fabsf(value_noise(oseed[o], px * freq, py * freq, pz * freq));
            norm += w;
            w *= 0.5f;
            freq *= 2.0f;
        }
        const float v = fminf(sum / norm, 1.0f) * 65535.0f + 0.5f;
        out[i] = (uint16_t)v;
    }
}

__global__ __launch_bounds__(256) void histogram_kernel(const uint16_t* __restrict__ vox, uint64_t n,
                                                        unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[kHistBins];
    for (int b = threadIdx.x; b < kHistBins; b += 256) h[b] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&h[vox[i] >> 4], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kHistBins; b += 256)
        if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

// v < thr -> 0; v in [thr, vmax] -> [1, 65535] (linear, clamped).
__global__ __launch_bounds__(256) void threshold_kernel(uint16_t* __restrict__ vox, uint64_t n, uint32_t thr,
                                                        uint32_t vmax, unsigned long long* __restrict__ zeros) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long z = 0;
    const uint32_t span = vmax > thr ? vmax - thr : 1u;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t v = vox[i];
        uint32_t o = 0;
        if (v >= thr) {
            const uint64_t r = 1u + ((uint64_t)(v - thr) * 65534u) / span;
            o = r < 65535u ? (uint32_t)r : 65535u;
        } else {
            ++z;
        }
        vox[i] = (uint16_t)o;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) z += __shfl_xor(z, off, 64);
    if ((threadIdx.x & 63) == 0 && z) atomicAdd(zeros, z);
}

static_assert(sizeof(scratch) == (kHistBins + 1) * sizeof(unsigned long long), "scratch size");

// Device address of `scratch` per device ordinal, set once per device by
// synth_scratch_init with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch[kMaxDevices] = {};

// Turbulence, histogram threshold and count, with the histogram (and the
// zero counter after it) in d_hist: kHistBins + 1 uint64.  clear_by_kernel:
// clear it with a fill kernel instead of hipMemsetAsync.
void turbulence_impl(hipdvr_ctx* ctx, uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, float empty_fraction,
                     uint16_t* d_out, double* empty_out, unsigned long long* d_hist, bool clear_by_kernel) {
    hipStream_t st = ctx_stream(ctx);
    const uint64_t n = (uint64_t)nx * ny * nz;
    const uint32_t m = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
    const uint64_t want = (n + 255) / 256;
    const uint32_t grid = (uint32_t)(want < (uint64_t)cus * 16 ? want : (uint64_t)cus * 16);
    const uint32_t s32 = (uint32_t)(seed ^ (seed >> 32));
    hipLaunchKernelGGL(turbulence_kernel, dim3(grid), dim3(256), 0, st, d_out, nx, ny, n, 1.0f / (float)m, s32);
    hip_check(hipGetLastError(), "turbulence kernel");
    if (clear_by_kernel) {
        launch_zero_u64(d_hist, kHistBins + 1, st);
        hip_check(hipGetLastError(), "histogram clear");
    } else {
        hip_check(hipMemsetAsync(d_hist, 0, (kHistBins + 1) * sizeof(unsigned long long), st), "memset");
    }
    hipLaunchKernelGGL(histogram_kernel, dim3(grid), dim3(256), 0, st, (const uint16_t*)d_out, n, d_hist);
    hip_check(hipGetLastError(), "histogram kernel");
    unsigned long long hist[kHistBins];
    hip_check(hipMemcpyAsync(hist, d_hist, kHistBins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
              "histogram download");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    // bin boundary b*16 whose count below is nearest the target
    const double target = (double)empty_fraction * (double)n;
    uint32_t best = 0;
    double best_err = target, below = 0.0;
    for (int b = 0; b < kHistBins; ++b) {
        below += (double)hist[b];
        const double err = std::fabs(below - target);
        if (err < best_err) {
            best_err = err;
            best = (uint32_t)(b + 1);
        }
    }
    const uint32_t thr = best >= kHistBins ? 65535u : best * 16u;
    uint32_t top = kHistBins - 1;
    while (top > 0 && hist[top] == 0) --top;
    const uint32_t vmax = top * 16u + 15u;   // upper end of the highest occupied bin
    hipLaunchKernelGGL(threshold_kernel, dim3(grid), dim3(256), 0, st, d_out, n, thr, vmax, d_hist + kHistBins);
    hip_check(hipGetLastError(), "threshold kernel");
    unsigned long long zeros = 0;
    hip_check(hipMemcpyAsync(&zeros, d_hist + kHistBins, sizeof zeros, hipMemcpyDeviceToHost, st), "count download");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (empty_out) *empty_out = (double)zeros / (double)n;
}

void check_turbulence_args(uint32_t nx, uint32_t ny, uint32_t nz, float empty_fraction, const uint16_t* d_out) {
    if (!d_out || nx == 0 || ny == 0 || nz == 0) fail(HIPDVR_ERR_INVALID, "bad turbulence arguments");
    if (!(empty_fraction >= 0.0f && empty_fraction < 1.0f)) fail(HIPDVR_ERR_INVALID, "empty_fraction must be in [0, 1)");
}

}  // namespace

void synth_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) fail(HIPDVR_ERR_INVALID, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail(HIPDVR_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[device] = (unsigned long long*)p;
}

}  // namespace hipdvr

using namespace hipdvr;

extern "C" int hipdvr_synth_turbulence_device(hipdvr_ctx* ctx, uint64_t seed, uint32_t nx, uint32_t ny,
                                              uint32_t nz, float empty_fraction, uint16_t* d_out,
                                              double* empty_out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        check_turbulence_args(nx, ny, nz, empty_fraction, d_out);
        ctx_set_device(ctx);
        unsigned long long* d_hist = nullptr;
        hip_check(hipMalloc((void**)&d_hist, (kHistBins + 1) * sizeof(unsigned long long)), "hipMalloc(histogram)");
        try {
            turbulence_impl(ctx, seed, nx, ny, nz, empty_fraction, d_out, empty_out, d_hist, false);
        } catch (...) {
            (void)hipFree(d_hist);
            throw;
        }
        (void)hipFree(d_hist);
    });
}

extern "C" int hipdvr_synth_turbulence_device_noalloc(hipdvr_ctx* ctx, uint64_t seed, uint32_t nx, uint32_t ny,
                                                      uint32_t nz, float empty_fraction, uint16_t* d_out,
                                                      double* empty_out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        check_turbulence_args(nx, ny, nz, empty_fraction, d_out);
        ctx_set_device(ctx);
        unsigned long long* d_hist = nullptr;
        {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (ctx->device >= 0 && ctx->device < kMaxDevices) d_hist = g_scratch[ctx->device];
        }
        if (!d_hist) fail(HIPDVR_ERR_HIP, "the device's scratch address is not set up (hipdvr_create sets it)");
        turbulence_impl(ctx, seed, nx, ny, nz, empty_fraction, d_out, empty_out, d_hist, true);
    });
}
