// volume.hip -- volume upload: linear copy, 4^3-block tiled copy, brick
// min/max, and the per-TF brick occupancy used for exact empty-space skipping.
#include <hip/hip_runtime.h>

#include <string>

#include "cell_layout.h"
#include "errors.h"
#include "march_common.h"
#include "volume.h"

namespace hipdvr {

namespace {

// One lane per tiled element: gather from the linear copy (clamped), write
// the tiled copy contiguously.
// (Grid-stride: the element count of an 8 GiB volume reaches 2^32, beyond
// what one dispatch addresses.)
template <typename T>
__global__ __launch_bounds__(256) void tile_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                   uint32_t X, uint32_t Y, uint32_t Z, uint32_t nbx,
                                                   uint32_t nby, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t in = (uint32_t)(i & 63);
        const uint64_t blk = i >> 6;
        const uint32_t bx = (uint32_t)(blk % nbx);
        const uint64_t r = blk / nbx;
        const uint32_t by = (uint32_t)(r % nby);
        const uint32_t bz = (uint32_t)(r / nby);
        uint32_t x = bx * 4 + (in & 3), y = by * 4 + ((in >> 2) & 3), z = bz * 4 + (in >> 4);
        x = x < X ? x : X - 1;
        y = y < Y ? y : Y - 1;
        z = z < Z ? z : Z - 1;
        dst[i] = src[((uint64_t)z * Y + y) * X + x];
    }
}

// Workgroups of a grid-stride launch: one per per_block items, at most 2^20.
inline unsigned stride_grid(uint64_t items, uint64_t per_block = 256) {
    const uint64_t want = (items + per_block - 1) / per_block;
    const uint64_t cap = 1ull << 20;
    return (unsigned)(want < 1 ? 1 : (want < cap ? want : cap));
}

// One workgroup per brick (grid-stride over the bricks): min/max over
// [b*B, b*B + B] per axis, clamped.
template <typename T>
__global__ __launch_bounds__(256) void minmax_kernel(const T* __restrict__ src, uint2* __restrict__ mm,
                                                     uint32_t X, uint32_t Y, uint32_t Z, uint32_t bl,
                                                     uint32_t nbx, uint32_t nby, uint64_t nbricks) {
    for (uint64_t b64 = blockIdx.x; b64 < nbricks; b64 += gridDim.x) {
        const uint32_t b = (uint32_t)b64;
        const uint32_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
        const uint32_t B1 = (1u << bl) + 1;
        const uint32_t n = B1 * B1 * B1;
        uint32_t lo = 0xffffffffu, hi = 0;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
            uint32_t x = (bx << bl) + k % B1;
            uint32_t y = (by << bl) + (k / B1) % B1;
            uint32_t z = (bz << bl) + k / (B1 * B1);
            x = x < X ? x : X - 1;
            y = y < Y ? y : Y - 1;
            z = z < Z ? z : Z - 1;
            const uint32_t v = src[((uint64_t)z * Y + y) * X + x];
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t l2 = __shfl_xor(lo, off, 64), h2 = __shfl_xor(hi, off, 64);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        __shared__ uint32_t s_lo[4], s_hi[4];
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            s_lo[w] = lo;
            s_hi[w] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
                lo = s_lo[i] < lo ? s_lo[i] : lo;
                hi = s_hi[i] > hi ? s_hi[i] : hi;
            }
            mm[b] = make_uint2(lo, hi);
        }
        __syncthreads();   // s_lo / s_hi are rewritten for the next brick
    }
}

// One lane per record: the 8 footprint voxels of cell (x, y, z), clamped.
// Records in the order of cell_layout.h; cells of the padding hold zeros.
template <typename T>
__global__ __launch_bounds__(256) void cells_kernel(const T* __restrict__ src, void* __restrict__ dst, uint32_t X,
                                                    uint32_t Y, uint32_t Z, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t dims[3] = {X, Y, Z};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        uint32_t c[3];
        cell_of_record(i, dims, c);
        const uint32_t x = c[0], y = c[1], z = c[2];
        uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (x < X && y < Y && z < Z) {
            const uint32_t x1 = x + 1 < X ? x + 1 : X - 1;
            const uint32_t y1 = y + 1 < Y ? y + 1 : Y - 1;
            const uint32_t z1 = z + 1 < Z ? z + 1 : Z - 1;
            const uint64_t XY = (uint64_t)X * Y;
            const uint64_t b0 = (uint64_t)z * XY, b1 = (uint64_t)z1 * XY;
            // trilinear order: llB lrB urB ulB llF lrF urF ulF
            v[0] = src[b0 + (uint64_t)y * X + x];
            v[1] = src[b0 + (uint64_t)y * X + x1];
            v[2] = src[b0 + (uint64_t)y1 * X + x1];
            v[3] = src[b0 + (uint64_t)y1 * X + x];
            v[4] = src[b1 + (uint64_t)y * X + x];
            v[5] = src[b1 + (uint64_t)y * X + x1];
            v[6] = src[b1 + (uint64_t)y1 * X + x1];
            v[7] = src[b1 + (uint64_t)y1 * X + x];
        }
        if (sizeof(T) == 2) {
            ((uint4*)dst)[i] = make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16),
                                          v[6] | (v[7] << 16));
        } else {
            ((uint2*)dst)[i] = make_uint2(v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24),
                                          v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24));
        }
    }
}

template <typename T>
void build_layouts(hipdvr_volume* v, hipStream_t st) {
    const uint32_t X = v->desc.dims[0], Y = v->desc.dims[1], Z = v->desc.dims[2];
    const uint64_t total = v->tiled_elems;
    hipLaunchKernelGGL(tile_kernel<T>, dim3(stride_grid(total)), dim3(256), 0, st,
                       (const T*)v->d_linear, (T*)v->d_tiled, X, Y, Z, v->nblocks[0], v->nblocks[1],
                       total);
    hip_check(hipGetLastError(), "tile kernel launch");
    hipLaunchKernelGGL(minmax_kernel<T>, dim3(stride_grid(v->brick_count, 1)), dim3(256), 0, st,
                       (const T*)v->d_linear, v->d_minmax, X, Y, Z, v->desc.brick_log2,
                       v->nbricks[0], v->nbricks[1], v->brick_count);
    hip_check(hipGetLastError(), "min/max kernel launch");
}

__device__ __forceinline__ float norm_raw(uint32_t v, int dtype) {
    return dtype == 0 ? VoxNorm<uint8_t>::apply(v) : VoxNorm<uint16_t>::apply(v);
}

__device__ __forceinline__ int map_index(float n, const OccArgs& a) {
    const float rw = (n * a.rwm_scale) + a.rwm_offset;
    const float x = rw_to_norm(rw, a.dom_lo, a.dom_hi);
    const int w1 = a.tf_width - 1;
    const int idx = iround_tgt(x * (float)w1);
    return idx < 0 ? 0 : (idx > w1 ? w1 : idx);
}

__global__ __launch_bounds__(256) void occupancy_kernel(const OccArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const uint2 mm = a.minmax[i];
    const float n0 = norm_raw(mm.x, a.dtype) - 1e-5f;
    const float n1 = norm_raw(mm.y, a.dtype) + 1e-5f;
    int i0 = map_index(n0, a), i1 = map_index(n1, a);
    if (i0 > i1) {
        const int t = i0;
        i0 = i1;
        i1 = t;
    }
    i0 = i0 > 0 ? i0 - 1 : 0;
    i1 = i1 < a.tf_width - 1 ? i1 + 1 : a.tf_width - 1;
    a.occ[i] = (a.alpha_prefix[i1 + 1] - a.alpha_prefix[i0]) != 0u ? 1 : 0;
}

__global__ __launch_bounds__(256) void occupancy_bits_kernel(const uint8_t* __restrict__ occ,
                                                             uint32_t* __restrict__ bits, uint64_t count) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= count) return;
    uint32_t word = 0;
    for (uint32_t k = 0; k < 32 && w * 32 + k < count; ++k) word |= (occ[w * 32 + k] != 0 ? 1u : 0u) << k;
    bits[w] = word;
}

}  // namespace

void launch_occupancy(const OccArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(occupancy_kernel, dim3((unsigned)((a.count + 255) / 256)), dim3(256), 0, st, a);
    const uint64_t words = (a.count + 31) / 32;
    hipLaunchKernelGGL(occupancy_bits_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, a.occ, a.bits,
                       a.count);
}

hipdvr_volume* volume_create(hipdvr_ctx* ctx, hipStream_t st, const hipdvr_volume_desc* d,
                             const void* voxels, bool on_device) {
    for (int i = 0; i < 3; ++i)
        if (d->dims[i] < 1 || d->dims[i] > (1u << 20))
            fail(HIPDVR_ERR_INVALID, "volume dimensions must be in [1, 2^20]");
    if (d->dtype != HIPDVR_U8 && d->dtype != HIPDVR_U16)
        fail(HIPDVR_ERR_INVALID, "dtype must be HIPDVR_U8 or HIPDVR_U16");
    uint32_t bl = d->brick_log2 ? d->brick_log2 : 3;
    if (bl < 2 || bl > 5) fail(HIPDVR_ERR_INVALID, "brick_log2 must be in [2, 5]");
    hipdvr_volume* v = new hipdvr_volume();
    v->ctx = ctx;
    v->tiled_elems = 64;
    v->brick_count = 1;
    for (int i = 0; i < 3; ++i) {
        v->desc.dims[i] = d->dims[i];
        v->desc.spacing[i] = d->spacing[i];
        v->desc.offset[i] = d->offset[i];
        v->nblocks[i] = (d->dims[i] + 3) / 4;
        v->nbricks[i] = (d->dims[i] + (1u << bl) - 1) >> bl;
        v->tiled_elems *= v->nblocks[i];
        v->brick_count *= v->nbricks[i];
    }
    v->desc.dtype = d->dtype;
    v->desc.rwm_scale = d->rwm_scale;
    v->desc.rwm_offset = d->rwm_offset;
    v->desc.brick_log2 = bl;
    const size_t vsz = d->dtype == HIPDVR_U8 ? 1 : 2;
    const size_t bytes = (size_t)d->dims[0] * d->dims[1] * d->dims[2] * vsz;
    try {
        hip_check(hipMalloc(&v->d_linear, bytes), "hipMalloc(volume)");
        hip_check(hipMalloc(&v->d_tiled, v->tiled_elems * vsz), "hipMalloc(tiled volume)");
        hip_check(hipMalloc((void**)&v->d_minmax, v->brick_count * sizeof(uint2)), "hipMalloc(minmax)");
        hip_check(hipMalloc((void**)&v->d_occ, v->brick_count), "hipMalloc(occupancy)");
        v->occ_words = (uint32_t)((v->brick_count + 31) / 32);
        hip_check(hipMalloc((void**)&v->d_occ_bits, (size_t)v->occ_words * 4), "hipMalloc(occupancy bits)");
        hip_check(hipMemcpyAsync(v->d_linear, voxels, bytes,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
                  "volume upload");
        if (d->dtype == HIPDVR_U8) build_layouts<uint8_t>(v, st);
        else build_layouts<uint16_t>(v, st);
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    } catch (...) {
        volume_destroy(v);
        throw;
    }
    return v;
}

namespace {
// One workgroup: minimum and maximum over the brick table (every voxel lies
// in some brick's range).
__global__ __launch_bounds__(1024) void minmax_reduce_kernel(const uint2* __restrict__ mm, uint64_t n,
                                                            uint32_t* __restrict__ out) {
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint2 v = mm[i];
        lo = v.x < lo ? v.x : lo;
        hi = v.y > hi ? v.y : hi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t l2 = __shfl_xor(lo, off, 64), h2 = __shfl_xor(hi, off, 64);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    __shared__ uint32_t s_lo[16], s_hi[16];
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            lo = s_lo[w] < lo ? s_lo[w] : lo;
            hi = s_hi[w] > hi ? s_hi[w] : hi;
        }
        out[0] = lo;
        out[1] = hi;
    }
}
}  // namespace

void volume_raw_minmax(const hipdvr_volume* v, hipStream_t st, uint32_t out[2]) {
    uint32_t* d = nullptr;
    hip_check(hipMalloc((void**)&d, 8), "hipMalloc(minmax)");
    hipLaunchKernelGGL(minmax_reduce_kernel, dim3(1), dim3(1024), 0, st, v->d_minmax, v->brick_count, d);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = e1 == hipSuccess ? hipMemcpyAsync(out, d, 8, hipMemcpyDeviceToHost, st) : e1;
    const hipError_t e3 = e2 == hipSuccess ? hipStreamSynchronize(st) : e2;
    (void)hipFree(d);
    hip_check(e3, "volume min/max");
}

void launch_brick_minmax(const hipdvr_volume* v, uint32_t bl, const uint32_t nb[3], uint2* out, hipStream_t st) {
    const uint32_t X = v->desc.dims[0], Y = v->desc.dims[1], Z = v->desc.dims[2];
    const uint64_t count = (uint64_t)nb[0] * nb[1] * nb[2];
    const dim3 grid(stride_grid(count, 1));
    if (v->desc.dtype == HIPDVR_U8)
        hipLaunchKernelGGL(minmax_kernel<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)v->d_linear, out, X, Y,
                           Z, bl, nb[0], nb[1], count);
    else
        hipLaunchKernelGGL(minmax_kernel<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)v->d_linear, out, X,
                           Y, Z, bl, nb[0], nb[1], count);
}

uint64_t cell_record_bytes(const hipdvr_volume* v) {
    return cell_record_count(v->desc.dims) * (v->desc.dtype == HIPDVR_U8 ? 8 : 16);
}

bool cells_allowed(const hipdvr_volume* v) { return cell_record_bytes(v) <= kMaxCellBytes; }

void ensure_cells(hipdvr_volume* v, hipStream_t st) {
    if (v->d_cells) return;
    const uint64_t bytes = cell_record_bytes(v);
    if (bytes > kMaxCellBytes) fail(HIPDVR_ERR_UNSUPPORTED, "cell records of this volume exceed the size limit");
    const uint64_t total = bytes / (v->desc.dtype == HIPDVR_U8 ? 8 : 16);
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc(cell records)");
    const dim3 grid(stride_grid(total));
    if (v->desc.dtype == HIPDVR_U8)
        hipLaunchKernelGGL(cells_kernel<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)v->d_linear, p,
                           v->desc.dims[0], v->desc.dims[1], v->desc.dims[2], total);
    else
        hipLaunchKernelGGL(cells_kernel<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)v->d_linear, p,
                           v->desc.dims[0], v->desc.dims[1], v->desc.dims[2], total);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(st);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        (void)hipFree(p);
        hip_check(e1 != hipSuccess ? e1 : e2, "cell record build");
    }
    v->d_cells = p;
    v->cell_bytes = bytes;
    v->cells_wide = bytes > 0xffffff00ull;
}

void volume_destroy(hipdvr_volume* v) {
    if (!v) return;
    if (v->d_cells) (void)hipFree(v->d_cells);
    if (v->d_linear) (void)hipFree(v->d_linear);
    if (v->d_tiled) (void)hipFree(v->d_tiled);
    if (v->d_minmax) (void)hipFree(v->d_minmax);
    if (v->d_occ) (void)hipFree(v->d_occ);
    if (v->d_occ_bits) (void)hipFree(v->d_occ_bits);
    if (v->d_grad) (void)hipFree(v->d_grad);
    delete v;
}

}  // namespace hipdvr
