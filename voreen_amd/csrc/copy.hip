// copy.hip -- plain streaming device copy (16-byte vectors, grid-stride).
// Used as the known-size reference when HBM counters are calibrated.
#include "kernels.h"

namespace hipdvr {

__global__ __launch_bounds__(256) void copy16_kernel(uint4* __restrict__ dst,
                                                     const uint4* __restrict__ src, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

void launch_copy16(void* dst, const void* src, size_t n, hipStream_t st) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)dst,
                       (const uint4*)src, n);
}

}  // namespace hipdvr
