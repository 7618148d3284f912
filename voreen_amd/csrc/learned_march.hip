// learned_march.hip -- learned stepping: brick statistics, network features,
// trace targets, the learned marcher, the sparse plain re-march judge.
//
// Bricks are 16^3 voxels (kLBrick); a row of the network is (brick, ray
// octant), row = brick*8 + octant, octant bit i set when dir[i] < 0.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "learned.h"
#include "learned_ray.h"
#include "march_cell.h"
#include "march_common.h"

namespace hipdvr {

// ---------------------------------------------------------------------------
// per-brick statistics: min, max, mean, variance of the normalized intensity
// over the brick plus a one-voxel upper apron (clamped)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void brick_stats_kernel(const T* __restrict__ src, float4* __restrict__ out,
                                                          uint32_t X, uint32_t Y, uint32_t Z, uint32_t nbx,
                                                          uint32_t nby) {
    const uint32_t b = blockIdx.x;
    const uint32_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const uint32_t B1 = kLBrick + 1, n = B1 * B1 * B1;
    float lo = 1e30f, hi = -1e30f, s = 0.f, s2 = 0.f;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        uint32_t x = bx * kLBrick + k % B1, y = by * kLBrick + (k / B1) % B1, z = bz * kLBrick + k / (B1 * B1);
        x = x < X ? x : X - 1;
        y = y < Y ? y : Y - 1;
        z = z < Z ? z : Z - 1;
        const float v = VoxNorm<T>::apply(src[((uint64_t)z * Y + y) * X + x]);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
        s += v;
        s2 += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
        s += __shfl_xor(s, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    __shared__ float sh[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[w][0] = lo;
        sh[w][1] = hi;
        sh[w][2] = s;
        sh[w][3] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; ++i) {
            lo = fminf(lo, sh[i][0]);
            hi = fmaxf(hi, sh[i][1]);
            s += sh[i][2];
            s2 += sh[i][3];
        }
        const float mean = s / (float)n;
        out[b] = make_float4(lo, hi, mean, fmaxf(s2 / (float)n - mean * mean, 0.f));
    }
}

void launch_brick_stats(const void* linear, int dtype, const uint32_t dims[3], const uint32_t nb[3],
                        float4* out, hipStream_t st) {
    const dim3 grid(nb[0] * nb[1] * nb[2]);
    if (dtype == 0)
        hipLaunchKernelGGL(brick_stats_kernel<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)linear, out,
                           dims[0], dims[1], dims[2], nb[0], nb[1]);
    else
        hipLaunchKernelGGL(brick_stats_kernel<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)linear,
                           out, dims[0], dims[1], dims[2], nb[0], nb[1]);
}

// ---------------------------------------------------------------------------
// features: one row per (brick, octant)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int map_idx(float n, const FeatureArgs& a) {
    const float rw = (n * a.rwm_scale) + a.rwm_offset;
    const float x = rw_to_norm(rw, a.dom_lo, a.dom_hi);
    const int w1 = a.tf_width - 1;
    const int idx = iround_tgt(x * (float)w1);
    return idx < 0 ? 0 : (idx > w1 ? w1 : idx);
}

__global__ __launch_bounds__(256) void features_kernel(const FeatureArgs a) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.rows) return;
    const uint32_t b = row >> 3, oct = row & 7;
    const float4 st = a.stats[b];
    int i0 = map_idx(st.x, a), i1 = map_idx(st.y, a);
    if (i0 > i1) {
        const int tmp = i0;
        i0 = i1;
        i1 = tmp;
    }
    const float cnt = (float)(i1 - i0 + 1);
    const float mean_a = (a.alpha_prefix[i1 + 1] - a.alpha_prefix[i0]) / cnt;
    const float frac_nz = (float)(a.nz_prefix[i1 + 1] - a.nz_prefix[i0]) / cnt;
    const float a_mean = a.alpha[map_idx(st.z, a)];
    if (oct == 0) {
        // exact: as the 8^3 occupancy (range widened by 1e-5, one index each side)
        int j0 = map_idx(st.x - 1e-5f, a), j1 = map_idx(st.y + 1e-5f, a);
        if (j0 > j1) {
            const int tmp = j0;
            j0 = j1;
            j1 = tmp;
        }
        j0 = j0 > 0 ? j0 - 1 : 0;
        j1 = j1 < a.tf_width - 1 ? j1 + 1 : a.tf_width - 1;
        a.occ16[b] = (a.nz_prefix[j1 + 1] - a.nz_prefix[j0]) != 0u ? 1 : 0;
    }
    float* x = a.x + (size_t)row * kLFeatures;
    x[0] = st.x;
    x[1] = st.y;
    x[2] = st.z;
    x[3] = sqrtf(st.w);
    x[4] = mean_a;
    x[5] = a_mean;
    x[6] = frac_nz;
    x[7] = 1.f - powf(1.f - fminf(mean_a, 0.999999f), 32.f);
    x[8] = (oct & 1) ? -1.f : 1.f;
    x[9] = (oct & 2) ? -1.f : 1.f;
    x[10] = (oct & 4) ? -1.f : 1.f;
    for (int i = 11; i < kLFeatures; ++i) x[i] = 0.f;
}

void launch_features(const FeatureArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(features_kernel, dim3((a.rows + 255) / 256), dim3(256), 0, st, a);
}

// network outputs -> stepping table: bit 7 = skip, bits 0..1 = log2 multiplier
__global__ void table_kernel(const float* __restrict__ y, uint32_t rows, uint8_t* __restrict__ table,
                             float skip_logit, int max_log2) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const float y0 = y[(size_t)row * 2], y1 = y[(size_t)row * 2 + 1];
    int m = (int)floorf(y1 + 0.5f);
    m = m < 0 ? 0 : (m > max_log2 ? max_log2 : m);
    table[row] = (uint8_t)(((y0 > skip_logit) ? 0x80 : 0) | m);
}

void launch_table(const float* y, uint32_t rows, uint8_t* table, float skip_logit, int max_log2,
                  hipStream_t st) {
    hipLaunchKernelGGL(table_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, y, rows, table, skip_logit,
                       max_log2);
}

__global__ void fill_table_kernel(uint8_t* table, uint32_t rows, uint8_t v) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < rows) table[row] = v;
}

void launch_fill_table(uint8_t* table, uint32_t rows, uint8_t v, hipStream_t st) {
    hipLaunchKernelGGL(fill_table_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, table, rows, v);
}

// bricks whose value range maps to zero alpha are skipped whatever the table
// says (exact), so the marcher reads one byte per brick
__global__ void merge_occ_kernel(uint8_t* table, const uint8_t* __restrict__ occ16, uint32_t rows) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < rows && !occ16[row >> 3]) table[row] = (uint8_t)(table[row] | 0x80);
}

void launch_merge_occ(uint8_t* table, const uint8_t* occ16, uint32_t rows, hipStream_t st) {
    hipLaunchKernelGGL(merge_occ_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, table, occ16, rows);
}

// rows where the table departs from the plain march: a step multiple above 1,
// or a skip of a brick that may hold alpha > 0 (count[0], cleared first)
__global__ void table_gain_kernel(const uint8_t* __restrict__ table, const uint8_t* __restrict__ occ16,
                                  uint32_t rows, unsigned int* __restrict__ count) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    bool gain = false;
    if (row < rows) {
        const uint8_t e = table[row];
        gain = (e & 0x80) ? occ16[row >> 3] != 0 : (e & 3) != 0;
    }
    const unsigned long long b = __ballot(gain);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (unsigned int)__popcll(b));
}

__global__ void zero_u32_kernel(unsigned int* p) { *p = 0u; }

void launch_table_gain(const uint8_t* table, const uint8_t* occ16, uint32_t rows, unsigned int* count,
                       hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, count);
    hipLaunchKernelGGL(table_gain_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, table, occ16, rows, count);
}

// ---------------------------------------------------------------------------
// shared march helpers (linear filter, tiled volume, TF classification)
// ---------------------------------------------------------------------------
struct LFetch {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t sy, sz, dims[3];
    int u16;
    __device__ __forceinline__ uint32_t ox(uint32_t x) const { return ((x >> 2) << 6) | (x & 3); }
    __device__ __forceinline__ uint32_t oy(uint32_t y) const { return (y >> 2) * sy + ((y & 3) << 2); }
    __device__ __forceinline__ uint32_t oz(uint32_t z) const { return (z >> 2) * sz + ((z & 3) << 4); }
    __device__ __forceinline__ float raw(uint32_t e) const {
        return u16 ? (float)__builtin_amdgcn_raw_buffer_load_b16(rsrc, e * 2u, 0, 0)
                   : (float)__builtin_amdgcn_raw_buffer_load_b8(rsrc, e, 0, 0);
    }
};

__device__ __forceinline__ void lfetch_init(LFetch& f, const LearnedArgs& a) {
    f.sy = ((a.m.dims[0] + 3) >> 2) << 6;
    f.sz = f.sy * ((a.m.dims[1] + 3) >> 2);
    f.u16 = a.u16;
    for (int i = 0; i < 3; ++i) f.dims[i] = a.m.dims[i];
    const uint32_t bytes = f.sz * ((a.m.dims[2] + 3) >> 2) * (a.u16 ? 2u : 1u);
    f.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.m.bricks, (short)0, (int)bytes, 0x00020000);
}

// Trilinear sample at clamped voxel position vp (linear filter; raw scale).
__device__ __forceinline__ float lsample(const LFetch& f, const float vp[3], Corners& c) {
    corners(vp, f.dims, c);
    const uint32_t x0 = f.ox(c.lo[0]), x1 = f.ox(c.lo[0] + c.d[0]);
    const uint32_t y0 = f.oy(c.lo[1]), y1 = f.oy(c.lo[1] + c.d[1]);
    const uint32_t z0 = f.oz(c.lo[2]), z1 = f.oz(c.lo[2] + c.d[2]);
    const float ix = 1.f - c.p[0], iy = 1.f - c.p[1], iz = 1.f - c.p[2];
    float s = f.raw(x0 + y0 + z0) * ix * iy * iz;
    s += f.raw(x1 + y0 + z0) * c.p[0] * iy * iz;
    s += f.raw(x1 + y1 + z0) * c.p[0] * c.p[1] * iz;
    s += f.raw(x0 + y1 + z0) * ix * c.p[1] * iz;
    s += f.raw(x0 + y0 + z1) * ix * iy * c.p[2];
    s += f.raw(x1 + y0 + z1) * c.p[0] * iy * c.p[2];
    s += f.raw(x1 + y1 + z1) * c.p[0] * c.p[1] * c.p[2];
    s += f.raw(x0 + y1 + z1) * ix * c.p[1] * c.p[2];
    return s * (f.u16 ? (1.0f / 65535.0f) : (1.0f / 255.0f));
}

// Gradient by plane bilinears (same scheme as the bricked kernel).
__device__ __forceinline__ void lgradient(const LFetch& f, const MarchArgs& a, const float u[3],
                                          const Corners& c, float g[3]) {
    uint32_t o[3][2];
    o[0][0] = f.ox(c.lo[0]);
    o[0][1] = f.ox(c.lo[0] + c.d[0]);
    o[1][0] = f.oy(c.lo[1]);
    o[1][1] = f.oy(c.lo[1] + c.d[1]);
    o[2][0] = f.oz(c.lo[2]);
    o[2][1] = f.oz(c.lo[2] + c.d[2]);
    float w[3][2];
    for (int i = 0; i < 3; ++i) {
        w[i][0] = 1.f - c.p[i];
        w[i][1] = c.p[i];
    }
    const float norm = f.u16 ? (1.0f / 65535.0f) : (1.0f / 255.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        const uint32_t m = a.dims[i] - 1;
        float pl[2][3];  // lo plane, hi plane, fraction of the two stencil points
        uint32_t P[4];
        for (int sgn = 0; sgn < 2; ++sgn) {
            const float uu = u[i] + (sgn ? 1.0f : -1.0f);
            const float cl = sminf(smaxf(uu, 0.0f), a.fdm1[i]);
            uint32_t lo = (uint32_t)cl;
            lo = lo < m ? lo : m;
            P[2 * sgn] = lo;
            P[2 * sgn + 1] = lo + 1 < m ? lo + 1 : m;
            pl[sgn][0] = cl - floorf(cl);
        }
        float B[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t oi = i == 0 ? f.ox(P[q]) : (i == 1 ? f.oy(P[q]) : f.oz(P[q]));
            float s = 0.f;
            for (int jj = 0; jj < 2; ++jj)
                for (int kk = 0; kk < 2; ++kk) s += f.raw(oi + o[j][jj] + o[k][kk]) * (w[j][jj] * w[k][kk]);
            B[q] = s;
        }
        const float vm = (1.f - pl[0][0]) * B[0] + pl[0][0] * B[1];
        const float vpp = (1.f - pl[1][0]) * B[2] + pl[1][0] * B[3];
        g[i] = (((vm - vpp) * norm) * a.grad_scale[i]) * a.rwm_scale;
    }
}


// ---------------------------------------------------------------------------
// trace targets: march each ray plainly (TF classification, no shading) and,
// for every 16^3 brick segment, compare the plain composite of the segment
// with the composites at 2x, 4x, 8x the step and with skipping it; keep the
// per-row maxima of the transmittance-weighted errors.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_pos(float* p, float v) {
    atomicMax((unsigned int*)p, __float_as_uint(fmaxf(v, 0.f)));
}

struct SegAcc {
    float r, g, b, a;
};

__device__ __forceinline__ void seg_add(SegAcc& s, float cr, float cg, float cb, float ca) {
    const float w = (1.f - s.a) * ca;
    s.r += w * cr;
    s.g += w * cg;
    s.b += w * cb;
    s.a += w;
}

__global__ __launch_bounds__(256) void trace_kernel(const LearnedArgs la, float* __restrict__ err,
                                                    unsigned int* __restrict__ visits) {
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[4][256];   // multipliers 1, 2, 4, 8
    const MarchArgs& a = la.m;
    for (int i = threadIdx.x; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    for (int k = 0; k < 4; ++k) s_alpha[k][threadIdx.x] = la.alpha_m[((1 << k) - 1) * 256 + threadIdx.x];
    __syncthreads();
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= a.w || py >= a.h) return;
    const size_t pix = (size_t)py * a.w + px;
    const float4 e = a.entry[pix], l = a.exit[pix];
    if (is_background(e, l)) return;
    LFetch f;
    lfetch_init(f, la);
    const float first[3] = {e.x, e.y, e.z}, last[3] = {l.x, l.y, l.z};
    float dir[3];
    const float tend = ray_setup(first, last, a.step, dir);
    const uint32_t oct = (dir[0] < 0.f ? 1u : 0u) | (dir[1] < 0.f ? 2u : 0u) | (dir[2] < 0.f ? 4u : 0u);
    const BoxExit bx = box_exit_setup(a, dir);
    float A = 0.f;
    float t = 0.f;
    int guard = 0;
    while (!(t > tend) && A < kErtOpacity && guard++ < 4096) {
        float s[3], u[3], vp[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = first[i] + t * dir[i];
            u[i] = (a.fdims[i] * s[i]) + (-0.5f);
            vp[i] = sminf(smaxf(u[i], 0.0f), a.fdm1[i]);
        }
        uint32_t bc[3];
        for (int i = 0; i < 3; ++i) bc[i] = (uint32_t)vp[i] >> kLBrickLog2;
        const uint32_t brick = (bc[2] * la.nb16[1] + bc[1]) * la.nb16[0] + bc[0];
        const float texit = fminf(kLearnedFastExit ? box_exit_fast(bx, first, dir, bc, kLBrick, la.nb16)
                                                   : box_exit(a, first, dir, bc, kLBrick, la.nb16),
                                  tend);
        // plain segment and the coarse variants from the same start
        SegAcc seg[4];
        for (int k = 0; k < 4; ++k) seg[k] = SegAcc{0.f, 0.f, 0.f, 0.f};
        float tt = t;
        int n = 0;
        // at least one sample per segment, so the ray always advances
        while (!(tt > tend) && (n == 0 || !(tt > texit)) && n < 4096) {
            float ss[3], vv[3];
            for (int i = 0; i < 3; ++i) {
                ss[i] = first[i] + tt * dir[i];
                vv[i] = sminf(smaxf((a.fdims[i] * ss[i]) + (-0.5f), 0.0f), a.fdm1[i]);
            }
            Corners c;
            const float inten = lsample(f, vv, c);
            const uint32_t c8 = s_tf[tf_index(inten, a)];
            const float cr = byte_norm(c8 & 0xffu), cg = byte_norm((c8 >> 8) & 0xffu),
                        cb = byte_norm((c8 >> 16) & 0xffu);
            const uint32_t ab = c8 >> 24;
            for (int k = 0; k < 4; ++k)
                if ((n & ((1 << k) - 1)) == 0 && ab) seg_add(seg[k], cr, cg, cb, s_alpha[k][ab]);
            tt += a.step;
            ++n;
        }
        const float trans = 1.f - A;
        const uint32_t row = brick * 8 + oct;
        atomicAdd(&visits[row], 1u);
        atomic_max_pos(&err[(size_t)row * 4 + 0], trans * seg[0].a);   // skipping loses this
        for (int k = 1; k < 4; ++k) {
            const float d = fmaxf(fmaxf(fabsf(seg[k].r - seg[0].r), fabsf(seg[k].g - seg[0].g)),
                                  fmaxf(fabsf(seg[k].b - seg[0].b), fabsf(seg[k].a - seg[0].a)));
            atomic_max_pos(&err[(size_t)row * 4 + k], trans * d);
        }
        A = A + (1.f - A) * seg[0].a;
        t = tt;
    }
}

void launch_trace(const LearnedArgs& la, float* err, unsigned int* visits, hipStream_t st) {
    const dim3 grid((la.m.w + 15) / 16, (la.m.h + 15) / 16);
    hipLaunchKernelGGL(trace_kernel, grid, dim3(256), 0, st, la, err, visits);
}

// errors -> training rows: targets (skip label, log2 multiplier label, weight)
__global__ void targets_kernel(const float* __restrict__ err, const unsigned int* __restrict__ visits,
                               const float* __restrict__ feat, uint32_t rows, float tau_skip, float tau_step,
                               float* __restrict__ out_x, float* __restrict__ out_t,
                               unsigned int* __restrict__ count, uint32_t capacity) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows || visits[row] == 0) return;
    const float* e = err + (size_t)row * 4;
    float m = 0.f;
    for (int k = 1; k < 4; ++k) {
        if (e[k] < tau_step) m = (float)k;
        else break;
    }
    const unsigned int slot = atomicAdd(count, 1u);
    if (slot >= capacity) return;
    for (int i = 0; i < kLFeatures; ++i) out_x[(size_t)slot * kLFeatures + i] = feat[(size_t)row * kLFeatures + i];
    out_t[(size_t)slot * 3 + 0] = e[0] < tau_skip ? 1.f : 0.f;
    out_t[(size_t)slot * 3 + 1] = m;
    out_t[(size_t)slot * 3 + 2] = e[0] < tau_skip ? 0.1f : 1.f;   // multiplier matters where kept
}

void launch_targets(const float* err, const unsigned int* visits, const float* feat, uint32_t rows,
                    float tau_skip, float tau_step, float* out_x, float* out_t, unsigned int* count,
                    uint32_t capacity, hipStream_t st) {
    hipLaunchKernelGGL(targets_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, err, visits, feat, rows,
                       tau_skip, tau_step, out_x, out_t, count, capacity);
}

template <typename T, bool WIDE, bool PHONG, bool COUNT>
__global__ __launch_bounds__(256) void learned_kernel(const LearnedArgs la) {
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];   // opacity correction of the base step
    const MarchArgs& a = la.m;
    for (int i = threadIdx.x; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[threadIdx.x] = la.alpha_m[threadIdx.x];
    __syncthreads();
    const uint32_t tiles_x = (a.w + 15) / 16;
    const uint32_t tiles_y = (a.h + 15) / 16;
    uint32_t tile_x, tile_y;   // XCD-compact tile order (xcd_tile)
    const bool has_tile = xcd_tile(blockIdx.x, tiles_x, tiles_y, tile_x, tile_y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = (int)tile_x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)tile_y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const CellFetch<T, WIDE> f = cell_fetch_of<T, WIDE>(a);
    unsigned long long n_samples = 0, n_shaded = 0;
    if (has_tile && px < a.w && py < a.h) {
        const size_t pix = (size_t)py * a.w + px;
        const float4 e = a.entry[pix], l = a.exit[pix];
        float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!is_background(e, l)) res = learned_ray<T, WIDE, PHONG>(la, f, s_tf, s_alpha, px, py, e, l, n_samples, n_shaded);
        a.out[pix] = res;
    }
    if (COUNT) {
        wave_add(a.counters + 1, n_samples);
        wave_add(a.counters + 3, n_shaded);
    }
}

template <typename T, bool WIDE>
static void launch_learned_t(const LearnedArgs& la, bool phong, bool count, dim3 grid, hipStream_t st) {
    if (phong) {
        if (count) hipLaunchKernelGGL((learned_kernel<T, WIDE, true, true>), grid, dim3(256), 0, st, la);
        else hipLaunchKernelGGL((learned_kernel<T, WIDE, true, false>), grid, dim3(256), 0, st, la);
    } else {
        const size_t shm = unshaded_dyn_lds(0);   // five workgroups per CU
        if (count) hipLaunchKernelGGL((learned_kernel<T, WIDE, false, true>), grid, dim3(256), shm, st, la);
        else hipLaunchKernelGGL((learned_kernel<T, WIDE, false, false>), grid, dim3(256), shm, st, la);
    }
}

void launch_learned(const LearnedArgs& la, bool phong, bool count, bool wide, hipStream_t st) {
    const dim3 grid(xcd_tile_grid((la.m.w + 15) / 16, (la.m.h + 15) / 16));
    if (la.u16) {
        if (wide) launch_learned_t<uint16_t, true>(la, phong, count, grid, st);
        else launch_learned_t<uint16_t, false>(la, phong, count, grid, st);
    } else {
        if (wide) launch_learned_t<uint8_t, true>(la, phong, count, grid, st);
        else launch_learned_t<uint8_t, false>(la, phong, count, grid, st);
    }
}

// ---------------------------------------------------------------------------
// judge: squared error of the learned image at the sparse pixels against the
// plainly re-marched sparse image; sets flag[0] = 1 when the PSNR is below
// the floor (mse > mse_max)
// ---------------------------------------------------------------------------
// One workgroup sums every sparse pixel's error (16K pixels at 1024^2, a few
// microseconds) and writes the sum and the flag: one launch, no atomics, the
// same sum on every run.
constexpr int kJudgeThreads = 1024;

__device__ __forceinline__ void judge_view(const float4* __restrict__ learned, int w, int h,
                                           const float4* __restrict__ sparse, int sw, int sh, int stride, int block,
                                           int offs, float denom, float mse_max, float* __restrict__ sse,
                                           int* __restrict__ flag) {
    __shared__ float part[kJudgeThreads / 64];
    float e = 0.f;
    for (int i = threadIdx.x; i < sw * sh; i += kJudgeThreads) {
        const int x = judge_pixel(i % sw, stride, block, offs), y = judge_pixel(i / sw, stride, block, offs);
        const float4 p = learned[(size_t)y * w + x], q = sparse[i];
        e += (p.x - q.x) * (p.x - q.x) + (p.y - q.y) * (p.y - q.y) + (p.z - q.z) * (p.z - q.z) +
             (p.w - q.w) * (p.w - q.w);
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int k = 0; k < kJudgeThreads / 64; ++k) t += part[k];
        sse[0] = t;
        flag[0] = (t / denom > mse_max) ? 1 : 0;
    }
}

__global__ __launch_bounds__(kJudgeThreads) void judge_kernel(const float4* __restrict__ learned, int w, int h,
                                                              const float4* __restrict__ sparse, int sw, int sh,
                                                              int stride, int block, int offs, float denom,
                                                              float mse_max, float* __restrict__ sse,
                                                              int* __restrict__ flag) {
    judge_view(learned, w, h, sparse, sw, sh, stride, block, offs, denom, mse_max, sse, flag);
}

// the judges of many views, one workgroup each (the same sums as judge_kernel)
__global__ __launch_bounds__(kJudgeThreads) void judge_multi_kernel(const JudgeJob* __restrict__ jobs) {
    const JudgeJob& j = jobs[blockIdx.x];
    judge_view(j.learned, j.w, j.h, j.sparse, j.sw, j.sh, j.stride, j.block, j.offs, j.denom, j.mse_max, j.sse,
               j.flag);
}

void launch_judge_multi(const JudgeJob* jobs, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(judge_multi_kernel, dim3(n), dim3(kJudgeThreads), 0, st, jobs);
}

void launch_judge(const float4* learned, int w, int h, const float4* sparse, int sw, int sh, int stride,
                  int block, int offs, float* sse, int* flag, float mse_max, hipStream_t st) {
    hipLaunchKernelGGL(judge_kernel, dim3(1), dim3(kJudgeThreads), 0, st, learned, w, h, sparse, sw, sh, stride,
                       block, offs, (float)(sw * sh * 4), mse_max, sse, flag);
}

}  // namespace hipdvr
