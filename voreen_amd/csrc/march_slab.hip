// march_slab.hip -- per-frame launches of the slab-staged ray march (device
// code and scheduling in march_slab.h): one workgroup of 256 rays per 16x16
// pixel tile, tiles in XCD-compact order.
#include <climits>

#include "kernels.h"
#include "march_common.h"
#include "march_slab.h"

namespace hipdvr {

namespace {

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Four consecutive voxels of a row as one word (8 bytes of uint16, 4 of uint8).
template <typename T>
struct Quad;
template <>
struct Quad<uint16_t> {
    typedef uint2 type;
};
template <>
struct Quad<uint8_t> {
    typedef uint32_t type;
};

template <typename T>
__device__ __forceinline__ typename Quad<T>::type pack4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    if constexpr (sizeof(T) == 2) {
        uint2 q;
        q.x = a | (b << 16);
        q.y = c | (d << 16);
        return q;
    } else {
        return a | (b << 8) | (c << 16) | (d << 24);
    }
}

// Lower-corner range, along axis i, of the ray's samples from t up to where
// its corner along the slab axis d leaves slab `slab`: [lo, hi] of
// floor(voxel position), widened by one voxel for the float error of the
// estimate.
__device__ __forceinline__ void segment_range(const MarchArgs& a, const SlabRay& ray, int d, int slab, int i,
                                              int& lo, int& hi) {
    float t1 = ray.tend;
    const float dd = ray.dir[d];
    if (dd > 0.0f) {
        const float vb = (float)((slab + 1) * kSlab);
        t1 = fminf(t1, ((vb + 0.5f) / a.fdims[d] - ray.first[d]) / dd);
    } else if (dd < 0.0f) {
        const float vb = (float)(slab * kSlab);
        t1 = fminf(t1, ((vb + 0.5f) / a.fdims[d] - ray.first[d]) / dd);
    }
    t1 = fmaxf(t1, ray.t);
    const float v0 = a.fdims[i] * (ray.first[i] + ray.t * ray.dir[i]) - 0.5f;
    const float v1 = a.fdims[i] * (ray.first[i] + t1 * ray.dir[i]) - 0.5f;
    lo = (int)floorf(fminf(v0, v1)) - 1;
    hi = (int)floorf(fmaxf(v0, v1)) + 1;
}

template <typename T, int SHADE, bool COUNT>
__global__ __launch_bounds__(256) void march_slab_kernel(const MarchArgs a, const uint32_t tiles_x,
                                                         const uint32_t tiles_y) {
    uint32_t tile_x, tile_y;
    if (!xcd_tile(blockIdx.x, tiles_x, tiles_y, tile_x, tile_y)) return;   // whole workgroup leaves together
    if (a.cond_flag && *a.cond_flag == 0) return;
    typedef typename Quad<T>::type quad_t;
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    __shared__ quad_t s_box4[kBoxCap<T> / 4];
    __shared__ float s_dir[4][3];
    __shared__ int s_key[2][4];
    __shared__ int s_bb[4][4];
    extern __shared__ uint32_t s_occ[];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    const bool occ_lds = a.occ_words <= kMaxLdsOccWords;
    if (occ_lds)
        for (uint32_t i = tid; i < a.occ_words; i += 256) s_occ[i] = a.occ_bits[i];

    const int lane = tid & 63, wave = tid >> 6;
    const int gx = (int)tile_x * 16 + (wave & 1) * 8 + (lane & 7);
    const int gy = (int)tile_y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int px = gx * a.stride + a.pix_offset, py = gy * a.stride + a.pix_offset;
    const bool inside = (gx < a.grid_w) && (gy < a.grid_h) && (px < a.w) && (py < a.h);
    RayCounters n;
    SlabRay ray;
    ray.active = false;
    ray.r = ray.g = ray.b = ray.A = 0.f;
    ray.t = 0.f;
    ray.loop = 0;
    ray.px = px;
    ray.py = py;
    ray.cur_brick = 0xffffffffu;
    ray.cur_occ = false;
    ray.t_lim = 0.f;
    ray.tend = 0.f;
    for (int i = 0; i < 3; ++i) ray.first[i] = ray.dir[i] = 0.f;
    if (inside) {
        const size_t pix = (size_t)py * a.w + px;
        const float4 e = a.entry[pix];
        const float4 l = a.exit[pix];
        if (!is_background(e, l)) {
            ++n.rays;
            ray.first[0] = e.x;
            ray.first[1] = e.y;
            ray.first[2] = e.z;
            const float last[3] = {l.x, l.y, l.z};
            ray.tend = ray_setup(ray.first, last, a.step, ray.dir);
            ray.active = true;
        }
    }
    // the slab axis: the largest component of the tile's summed ray direction
    {
        const float sx = wave_sum_f(ray.dir[0]), sy = wave_sum_f(ray.dir[1]), sz = wave_sum_f(ray.dir[2]);
        if (lane == 0) {
            s_dir[wave][0] = sx;
            s_dir[wave][1] = sy;
            s_dir[wave][2] = sz;
        }
    }
    __syncthreads();
    float dsum[3];
    for (int i = 0; i < 3; ++i) dsum[i] = s_dir[0][i] + s_dir[1][i] + s_dir[2][i] + s_dir[3][i];
    int d = 0;
    if (fabsf(dsum[1]) > fabsf(dsum[d])) d = 1;
    if (fabsf(dsum[2]) > fabsf(dsum[d])) d = 2;
    const int sgn = dsum[d] < 0.0f ? -1 : 1;
    const int ua = d == 0 ? 1 : 0, va = d == 2 ? 1 : 2;   // the two other axes

    SlabReader<T> rd;
    rd.box = (const T*)s_box4;
    rd.vol = (const T*)a.vox;
    for (int i = 0; i < 3; ++i) {
        rd.dims[i] = a.dims[i];
        rd.g.b0[i] = 0;
        rd.g.e[i] = 0;
    }
    rd.g.sy = rd.g.sz = 0;
    const uint32_t* occ = occ_lds ? s_occ : nullptr;

    // cross the leading empty bricks; key = the first slab needing data
    // (signed so that the tile's direction is increasing), kNoSlab if none
    int key = kNoSlab;
    if (ray.active) {
        const int sl = slab_march<T, SHADE, COUNT>(a, rd, s_tf, s_alpha, occ, d, -1, ray, n);
        key = sl == kNoSlab ? kNoSlab : sgn * sl;
    }

    for (int round = 0;; ++round) {
        const int par = round & 1;
        {
            const int k = wave_min_i(key);
            if (lane == 0) s_key[par][wave] = k;
        }
        __syncthreads();
        int kmin = s_key[par][0];
        for (int w = 1; w < 4; ++w) kmin = s_key[par][w] < kmin ? s_key[par][w] : kmin;
        if (kmin == kNoSlab) break;
        const bool last_round = round >= kMaxSlabRounds;
        const int slab = sgn * kmin;
        const bool mine = ray.active && (last_round || key == kmin);
        // lower-corner bounds of the segments of this round's rays
        int b[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
        if (mine && !last_round) {
            segment_range(a, ray, d, slab, ua, b[0], b[1]);
            segment_range(a, ray, d, slab, va, b[2], b[3]);
        }
        b[0] = wave_min_i(b[0]);
        b[1] = wave_max_i(b[1]);
        b[2] = wave_min_i(b[2]);
        b[3] = wave_max_i(b[3]);
        if (lane == 0)
            for (int k = 0; k < 4; ++k) s_bb[wave][k] = b[k];
        __syncthreads();
        int bu0 = s_bb[0][0], bu1 = s_bb[0][1], bv0 = s_bb[0][2], bv1 = s_bb[0][3];
        for (int w = 1; w < 4; ++w) {
            bu0 = s_bb[w][0] < bu0 ? s_bb[w][0] : bu0;
            bu1 = s_bb[w][1] > bu1 ? s_bb[w][1] : bu1;
            bv0 = s_bb[w][2] < bv0 ? s_bb[w][2] : bv0;
            bv1 = s_bb[w][3] > bv1 ? s_bb[w][3] : bv1;
        }
        // the box: corners [b - 1, b + 2] on every axis (footprint and the
        // gradient's outer planes), in virtual coordinates [0, dim + 1]
        int lo3[3], hi3[3];
        lo3[d] = slab * kSlab - 1;
        hi3[d] = slab * kSlab + kSlab + 1;
        lo3[ua] = bu0 - 1;
        hi3[ua] = bu1 + 2;
        lo3[va] = bv0 - 1;
        hi3[va] = bv1 + 2;
        int e3[3];
        bool staged = !last_round && bu0 <= bu1 && bv0 <= bv1;
        if (staged) {
            for (int i = 0; i < 3; ++i) {
                lo3[i] = clampi(lo3[i], 0, (int)a.dims[i] + 1);
                hi3[i] = clampi(hi3[i], 0, (int)a.dims[i] + 1);
            }
            lo3[0] &= ~3;                                  // x rows in aligned quads
            for (int i = 0; i < 3; ++i) e3[i] = hi3[i] - lo3[i] + 1;
            e3[0] = (e3[0] + 3) & ~3;
            // clip to the LDS capacity around the centre of the two cross axes
            auto vox = [&] { return (long long)e3[0] * e3[1] * e3[2]; };
            for (int it = 0; it < 64 && vox() > kBoxCap<T>; ++it) {
                const int c = e3[ua] >= e3[va] ? ua : va;
                const int cut = c == 0 ? 4 : 2;
                if (e3[c] <= cut + 4) break;
                e3[c] -= cut;
                lo3[c] += c == 0 ? (((it >> 1) & 1) ? 4 : 0) : 1;
            }
            staged = vox() <= kBoxCap<T>;
        }
        // stage: quads of 4 voxels along x, rows (y, z) of the box
        if (staged) {
            const int q = e3[0] >> 2, nslots = q * e3[1] * e3[2];
            const float inv_q = 1.0f / (float)q, inv_e1 = 1.0f / (float)e3[1];
            const int X = (int)a.dims[0], mx = X - 1, my = (int)a.dims[1] - 1, mz = (int)a.dims[2] - 1;
            const T* vol = (const T*)a.vox;
            const bool aligned_rows = (X & 3) == 0;
            for (int sl = tid; sl < nslots; sl += 256) {
                // quotients by a float reciprocal, corrected by one step
                // (the estimate is off by at most one)
                int row = (int)(((float)sl + 0.5f) * inv_q);
                int xq = sl - row * q;
                if (xq < 0) { --row; xq += q; } else if (xq >= q) { ++row; xq -= q; }
                int zr = (int)(((float)row + 0.5f) * inv_e1);
                int yr = row - zr * e3[1];
                if (yr < 0) { --zr; yr += e3[1]; } else if (yr >= e3[1]) { ++zr; yr -= e3[1]; }
                const int y = clampi(lo3[1] + yr, 0, my), z = clampi(lo3[2] + zr, 0, mz);
                const int x0 = lo3[0] + 4 * xq;
                const size_t base = ((size_t)z * a.dims[1] + (size_t)y) * (size_t)X;
                quad_t w;
                if (aligned_rows && x0 + 3 <= mx) {
                    w = *(const quad_t*)(vol + base + (size_t)x0);
                } else {
                    w = pack4<T>(vol[base + (size_t)clampi(x0, 0, mx)], vol[base + (size_t)clampi(x0 + 1, 0, mx)],
                                 vol[base + (size_t)clampi(x0 + 2, 0, mx)], vol[base + (size_t)clampi(x0 + 3, 0, mx)]);
                }
                s_box4[sl] = w;
            }
            for (int i = 0; i < 3; ++i) {
                rd.g.b0[i] = lo3[i];
                rd.g.e[i] = e3[i];
            }
            rd.g.sy = e3[0];
            rd.g.sz = e3[0] * e3[1];
        } else {
            for (int i = 0; i < 3; ++i) rd.g.e[i] = 0;
        }
        __syncthreads();
        if (mine) {
            const int sl = slab_march<T, SHADE, COUNT>(a, rd, s_tf, s_alpha, occ, d, last_round ? kAnySlab : slab,
                                                       ray, n);
            key = sl == kNoSlab ? kNoSlab : sgn * sl;
        }
        __syncthreads();   // the box is overwritten by the next round
    }

    if (inside) a.out[(size_t)gy * a.grid_w + gx] = make_float4(ray.r, ray.g, ray.b, ray.A);
    if (COUNT) {
        wave_add(a.counters + 0, n.rays);
        wave_add(a.counters + 1, n.samples);
        wave_add(a.counters + 2, n.fetched);
        wave_add(a.counters + 3, n.shaded);
        wave_add_max64(a.counters + 4, n.slots);
        wave_add_max64(a.counters + 5, n.shade_slots);
    }
}

template <typename T, int SHADE>
void launch_t(const MarchArgs& a, bool count, hipStream_t st) {
    const uint32_t tx = (a.grid_w + 15) / 16, ty = (a.grid_h + 15) / 16;
    const dim3 grid(xcd_tile_grid(tx, ty));
    const size_t shm = a.occ_words <= kMaxLdsOccWords ? (size_t)a.occ_words * 4 : 0;
    if (count)
        hipLaunchKernelGGL((march_slab_kernel<T, SHADE, true>), grid, dim3(256), shm, st, a, tx, ty);
    else
        hipLaunchKernelGGL((march_slab_kernel<T, SHADE, false>), grid, dim3(256), shm, st, a, tx, ty);
}

template <typename T>
void launch_s(const MarchArgs& a, int shade, bool count, hipStream_t st) {
    if (shade == 2) launch_t<T, 2>(a, count, st);
    else if (shade == 1) launch_t<T, 1>(a, count, st);
    else launch_t<T, 0>(a, count, st);
}

}  // namespace

void launch_march_slab(const MarchArgs& a, int dtype, int shade, bool count, hipStream_t st) {
    if (dtype == 0) launch_s<uint8_t>(a, shade, count, st);
    else launch_s<uint16_t>(a, shade, count, st);
}

}  // namespace hipdvr
