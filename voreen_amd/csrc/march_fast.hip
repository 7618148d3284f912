// march_fast.hip -- ray march on the 4^3-block tiled volume with exact
// empty-space skipping.
//
// Same sample positions, reconstruction, classification and compositing as
// march_ref.hip (and as CPURaycaster::directRendering,
// modules/base/processors/render/cpuraycaster.cpp:210-326), so the unshaded
// image is bit-identical.  Differences are only in where the voxels come
// from and which samples are reconstructed:
//   * voxels are read from the 4x4x4-block tiled copy (separable addressing,
//     one 128-byte line per uint16 block) with 32-bit-offset buffer loads;
//   * a sample whose footprint lies in a brick the occupancy table marks empty
//     maps to a zero-alpha TF entry, so it is not reconstructed; inside such a
//     brick the ray fast-forwards with t += step (the reference's float
//     sequence) up to an analytic bound of the brick exit;
//   * tiles run in XCD-compact order (xcd_tile), so the rays in flight on
//     all XCDs cover one compact image region and share the bricks in the
//     Infinity Cache, and each XCD's rays share its L2;
//   * Phong's central differences (mod_gradients.frag:202-214: six trilinear
//     lookups at +-1 voxel) are evaluated from per-axis plane bilinears that
//     share the sample's weights: the same clamped stencil positions and
//     interpolation, about a fifth of the arithmetic, equal to the six-lookup
//     form up to float rounding.
#include "kernels.h"
#include "march_common.h"

namespace hipdvr {

// Voxel source with separable block addressing.  WIDE volumes (more than
// 2^31 elements) use 64-bit global loads; the rest use buffer loads with a
// 32-bit byte offset from one descriptor.
template <typename T, bool WIDE>
struct TiledFetch {
    typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type off_t;
    const T* p;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t sy, sz;   // elements per block row (64*NBX) and per block slab (64*NBX*NBY)

    __device__ __forceinline__ off_t ox(uint32_t x) const { return (off_t)(((x >> 2) << 6) | (x & 3)); }
    __device__ __forceinline__ off_t oy(uint32_t y) const { return (off_t)(y >> 2) * sy + ((y & 3) << 2); }
    __device__ __forceinline__ off_t oz(uint32_t z) const { return (off_t)(z >> 2) * sz + ((z & 3) << 4); }

    __device__ __forceinline__ uint32_t load(off_t e) const {
        if (WIDE) return p[e];
        if (sizeof(T) == 2)
            return __builtin_amdgcn_raw_buffer_load_b16(rsrc, (uint32_t)e * 2u, 0, 0);
        return __builtin_amdgcn_raw_buffer_load_b8(rsrc, (uint32_t)e, 0, 0);
    }

    __device__ __forceinline__ void fetch8(const Corners& c, float v[8]) const {
        const off_t x0 = ox(c.lo[0]), x1 = ox(c.lo[0] + c.d[0]);
        const off_t y0 = oy(c.lo[1]), y1 = oy(c.lo[1] + c.d[1]);
        const off_t z0 = oz(c.lo[2]), z1 = oz(c.lo[2] + c.d[2]);
        const uint32_t r0 = load(x0 + y0 + z0);
        const uint32_t r1 = load(x1 + y0 + z0);
        const uint32_t r2 = load(x1 + y1 + z0);
        const uint32_t r3 = load(x0 + y1 + z0);
        const uint32_t r4 = load(x0 + y0 + z1);
        const uint32_t r5 = load(x1 + y0 + z1);
        const uint32_t r6 = load(x1 + y1 + z1);
        const uint32_t r7 = load(x0 + y1 + z1);
        v[0] = VoxNorm<T>::apply(r0);
        v[1] = VoxNorm<T>::apply(r1);
        v[2] = VoxNorm<T>::apply(r2);
        v[3] = VoxNorm<T>::apply(r3);
        v[4] = VoxNorm<T>::apply(r4);
        v[5] = VoxNorm<T>::apply(r5);
        v[6] = VoxNorm<T>::apply(r6);
        v[7] = VoxNorm<T>::apply(r7);
    }
    __device__ __forceinline__ float nearest(const float vp[3]) const {
        const uint32_t x = (uint32_t)iround_tgt(vp[0]);
        const uint32_t y = (uint32_t)iround_tgt(vp[1]);
        const uint32_t z = (uint32_t)iround_tgt(vp[2]);
        return VoxNorm<T>::apply(load(ox(x) + oy(y) + oz(z)));
    }
};

// Stencil coordinate along one axis: clamp(u +- 1) split into the lower plane,
// the upper plane (clamped) and the fraction.
struct Plane2 {
    uint32_t lo, hi;
    float p;
};

__device__ __forceinline__ Plane2 stencil_plane(float u, float fdm1, uint32_t m) {
    const float c = sminf(smaxf(u, 0.0f), fdm1);
    Plane2 r;
    r.p = c - floorf(c);
    r.lo = (uint32_t)c;
    r.lo = r.lo < m ? r.lo : m;
    r.hi = r.lo + 1 < m ? r.lo + 1 : m;
    return r;
}

// Central-difference gradient of the trilinear interpolant at the voxel
// position (u unclamped, c = corners of the clamped position): for axis i the
// two stencil samples clamp(u_i -+ 1) are interpolated along i between plane
// bilinears over the other two axes, which use the sample's own (clamped)
// coordinates and weights.  Raw voxel values; scaled once at the end.
template <typename T, bool WIDE>
__device__ __forceinline__ void gradient_planes(const TiledFetch<T, WIDE>& f, const MarchArgs& a,
                                                const float u[3], const Corners& c, float g[3]) {
    typedef typename TiledFetch<T, WIDE>::off_t off_t;
    off_t o[3][2];
    o[0][0] = f.ox(c.lo[0]);
    o[0][1] = f.ox(c.lo[0] + c.d[0]);
    o[1][0] = f.oy(c.lo[1]);
    o[1][1] = f.oy(c.lo[1] + c.d[1]);
    o[2][0] = f.oz(c.lo[2]);
    o[2][1] = f.oz(c.lo[2] + c.d[2]);
    float w[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        w[i][0] = 1.f - c.p[i];
        w[i][1] = c.p[i];
    }
    const float norm = sizeof(T) == 2 ? (1.0f / 65535.0f) : (1.0f / 255.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        const uint32_t m = a.dims[i] - 1;
        const Plane2 lo = stencil_plane(u[i] - 1.0f, a.fdm1[i], m);
        const Plane2 hi = stencil_plane(u[i] + 1.0f, a.fdm1[i], m);
        const uint32_t planes[4] = {lo.lo, lo.hi, hi.lo, hi.hi};
        float B[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const off_t oi = i == 0 ? f.ox(planes[q]) : (i == 1 ? f.oy(planes[q]) : f.oz(planes[q]));
            float s = 0.f;
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int kk = 0; kk < 2; ++kk)
                    s = s + (float)f.load(oi + o[j][jj] + o[k][kk]) * (w[j][jj] * w[k][kk]);
            B[q] = s;
        }
        const float vm = (1.f - lo.p) * B[0] + lo.p * B[1];
        const float vpl = (1.f - hi.p) * B[2] + hi.p * B[3];
        g[i] = (((vm - vpl) * norm) * a.grad_scale[i]) * a.rwm_scale;
    }
}

// Interior form of gradient_planes: when every axis has u_i in [1, dim_i - 3)
// and a proper fraction (the center footprint spans lo..lo+1), u_i +- 1 is
// exact in float, the stencil planes are lo-1, lo, lo+1, lo+2 with the
// sample's own fraction, and the two inner planes are the sample's corner
// values v[] (already normalized); only the outer planes are loaded.
template <typename T, bool WIDE>
__device__ __forceinline__ bool gradient_interior(const TiledFetch<T, WIDE>& f, const MarchArgs& a,
                                                  const float u[3], const Corners& c, const float v[8],
                                                  float g[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (!(u[i] >= 1.0f) || !(u[i] + 3.0f < a.fdims[i]) || c.d[i] == 0) return false;
    typedef typename TiledFetch<T, WIDE>::off_t off_t;
    const float norm = sizeof(T) == 2 ? (1.0f / 65535.0f) : (1.0f / 255.0f);
    const float px = c.p[0], py = c.p[1], pz = c.p[2];
    const float qx = 1.f - px, qy = 1.f - py, qz = 1.f - pz;
    // corner order: 0 (x0y0z0) 1 (x1y0z0) 2 (x1y1z0) 3 (x0y1z0) 4 (x0y0z1) 5 (x1y0z1) 6 (x1y1z1) 7 (x0y1z1)
    const off_t ox0 = f.ox(c.lo[0]), ox1 = f.ox(c.lo[0] + 1);
    const off_t oy0 = f.oy(c.lo[1]), oy1 = f.oy(c.lo[1] + 1);
    const off_t oz0 = f.oz(c.lo[2]), oz1 = f.oz(c.lo[2] + 1);
    // x: planes lo-1 and lo+2, bilinear over (y, z)
    {
        const off_t xm = f.ox(c.lo[0] - 1), xp = f.ox(c.lo[0] + 2);
        const float wyz[4] = {qy * qz, py * qz, qy * pz, py * pz};
        const off_t oyz[4] = {oy0 + oz0, oy1 + oz0, oy0 + oz1, oy1 + oz1};
        float bm = 0.f, bp = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bm = bm + (float)f.load(xm + oyz[k]) * wyz[k];
            bp = bp + (float)f.load(xp + oyz[k]) * wyz[k];
        }
        const float b0 = v[0] * wyz[0] + v[3] * wyz[1] + v[4] * wyz[2] + v[7] * wyz[3];
        const float b1 = v[1] * wyz[0] + v[2] * wyz[1] + v[5] * wyz[2] + v[6] * wyz[3];
        const float vm = qx * (bm * norm) + px * b0, vpl = qx * b1 + px * (bp * norm);
        g[0] = ((vm - vpl) * a.grad_scale[0]) * a.rwm_scale;
    }
    // y: planes lo-1 and lo+2, bilinear over (x, z)
    {
        const off_t ym = f.oy(c.lo[1] - 1), yp = f.oy(c.lo[1] + 2);
        const float wxz[4] = {qx * qz, px * qz, qx * pz, px * pz};
        const off_t oxz[4] = {ox0 + oz0, ox1 + oz0, ox0 + oz1, ox1 + oz1};
        float bm = 0.f, bp = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bm = bm + (float)f.load(ym + oxz[k]) * wxz[k];
            bp = bp + (float)f.load(yp + oxz[k]) * wxz[k];
        }
        const float b0 = v[0] * wxz[0] + v[1] * wxz[1] + v[4] * wxz[2] + v[5] * wxz[3];
        const float b1 = v[3] * wxz[0] + v[2] * wxz[1] + v[7] * wxz[2] + v[6] * wxz[3];
        const float vm = qy * (bm * norm) + py * b0, vpl = qy * b1 + py * (bp * norm);
        g[1] = ((vm - vpl) * a.grad_scale[1]) * a.rwm_scale;
    }
    // z: planes lo-1 and lo+2, bilinear over (x, y)
    {
        const off_t zm = f.oz(c.lo[2] - 1), zp = f.oz(c.lo[2] + 2);
        const float wxy[4] = {qx * qy, px * qy, qx * py, px * py};
        const off_t oxy[4] = {ox0 + oy0, ox1 + oy0, ox0 + oy1, ox1 + oy1};
        float bm = 0.f, bp = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bm = bm + (float)f.load(zm + oxy[k]) * wxy[k];
            bp = bp + (float)f.load(zp + oxy[k]) * wxy[k];
        }
        const float b0 = v[0] * wxy[0] + v[1] * wxy[1] + v[3] * wxy[2] + v[2] * wxy[3];
        const float b1 = v[4] * wxy[0] + v[5] * wxy[1] + v[7] * wxy[2] + v[6] * wxy[3];
        const float vm = qz * (bm * norm) + pz * b0, vpl = qz * b1 + pz * (bp * norm);
        g[2] = ((vm - vpl) * a.grad_scale[2]) * a.rwm_scale;
    }
    return true;
}

template <typename T, bool WIDE, int FILTER, bool PHONG, bool COUNT>
__global__ __launch_bounds__(256) void march_fast_kernel(const MarchArgs a, const uint32_t tiles_x,
                                                         const uint32_t tiles_y) {
    // XCD-compact tile order (xcd_tile)
    uint32_t tile_x, tile_y;
    if (!xcd_tile(blockIdx.x, tiles_x, tiles_y, tile_x, tile_y)) return;   // whole workgroup leaves together
    if (a.cond_flag && *a.cond_flag == 0) return;
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int gx = (int)tile_x * 16 + (wave & 1) * 8 + (lane & 7);
    const int gy = (int)tile_y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int px = gx * a.stride + a.pix_offset, py = gy * a.stride + a.pix_offset;
    const bool inside = (gx < a.grid_w) && (gy < a.grid_h) && (px < a.w) && (py < a.h);

    TiledFetch<T, WIDE> f;
    f.p = (const T*)a.bricks;
    f.sy = (uint32_t)((a.dims[0] + 3) >> 2) << 6;
    f.sz = f.sy * ((a.dims[1] + 3) >> 2);
    if (!WIDE) {
        const uint32_t bytes = (uint32_t)(f.sz * ((a.dims[2] + 3) >> 2) * sizeof(T));
        f.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.bricks, (short)0, (int)bytes, 0x00020000);
    }
    const uint32_t bl = a.brick_log2;
    const uint32_t nbx = a.nbricks[0], nbxy = a.nbricks[0] * a.nbricks[1];

    unsigned long long n_slots = 0, n_shade_slots = 0, n_samples = 0, n_fetched = 0, n_shaded = 0,
                       n_rays = 0;
    if (inside) {
        const size_t pix = (size_t)py * a.w + px;
        const float4 e = a.entry[pix];
        const float4 l = a.exit[pix];
        float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!is_background(e, l)) {
            n_rays = 1;
            const float first[3] = {e.x, e.y, e.z};
            const float last[3] = {l.x, l.y, l.z};
            float dir[3];
            const float tend = ray_setup(first, last, a.step, dir);
            const int cap = a.loop_cap;
            float t = 0.0f, r = 0.f, g = 0.f, bb = 0.f, A = 0.f;
            uint32_t cur_brick = 0xffffffffu;
            bool cur_occ = false;
            float t_lim = 0.0f;
            bool finished = false;
            for (int loop = 0; !finished && loop < cap; ++loop) {
                ++n_samples;
                bool shade_now = false;
                const float s[3] = {first[0] + t * dir[0], first[1] + t * dir[1],
                                    first[2] + t * dir[2]};
                float u[3], vp[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    u[i] = (a.fdims[i] * s[i]) + (-0.5f);
                    vp[i] = sminf(smaxf(u[i], 0.0f), a.fdm1[i]);
                }
                const uint32_t brick = (((uint32_t)vp[2] >> bl) * nbxy) +
                                       (((uint32_t)vp[1] >> bl) * nbx) + ((uint32_t)vp[0] >> bl);
                if (brick != cur_brick) {
                    cur_brick = brick;
                    cur_occ = a.brick_occ[brick] != 0;
                    if (!cur_occ) t_lim = empty_brick_limit(a, first, dir, vp, bl);
                }
                if (!cur_occ) {
                    // Fast-forward: every sample at t' < t_lim lies in this empty
                    // brick, so its only effect in the reference loop is the step
                    // count and t += step (same float sequence, same cap, same
                    // termination test).
                    float tn = t + a.step;
                    while (tn < t_lim && !(tn > tend) && loop + 1 < cap) {
                        t = tn;
                        ++loop;
                        ++n_samples;
                        if (COUNT) ++n_slots;
                        tn = t + a.step;
                    }
                } else {
                    // Two samples per pass: the gathers of the next sample (same
                    // brick, within the ray and the cap) are issued before this
                    // sample is composited, so their latency overlaps.  If this
                    // sample terminates the ray the next one is dropped unseen.
                    const float t1 = t + a.step;
                    float s1[3], u1[3], vp1[3];
                    bool pair = !(t1 > tend) && (loop + 1 < cap);
                    if (pair) {
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            s1[i] = first[i] + t1 * dir[i];
                            u1[i] = (a.fdims[i] * s1[i]) + (-0.5f);
                            vp1[i] = sminf(smaxf(u1[i], 0.0f), a.fdm1[i]);
                        }
                        const uint32_t brick1 = (((uint32_t)vp1[2] >> bl) * nbxy) +
                                                (((uint32_t)vp1[1] >> bl) * nbx) + ((uint32_t)vp1[0] >> bl);
                        pair = brick1 == cur_brick;
                    }
                    Corners c0, c1;
                    float v0[8], v1[8];
                    if (FILTER == 0) {
                        v0[0] = f.nearest(vp);
                        if (pair) v1[0] = f.nearest(vp1);
                    } else {
                        corners(vp, a.dims, c0);
                        f.fetch8(c0, v0);
                        if (pair) {
                            corners(vp1, a.dims, c1);
                            f.fetch8(c1, v1);
                        }
                    }
                    auto composite = [&](const float* vv, Corners& cc, const float* ss, const float* uu,
                                         const float* pp, float tt) {
                        ++n_fetched;
                        const float inten = FILTER == 0 ? vv[0] : trilerp(vv, cc.p);
                        const uint32_t c8 = s_tf[tf_index(inten, a)];
                        const float ca = s_alpha[c8 >> 24];
                        if (ca > 0.0f) {
                            shade_now = PHONG;
                            float cr = byte_norm(c8 & 0xffu);
                            float cg = byte_norm((c8 >> 8) & 0xffu);
                            float cb = byte_norm((c8 >> 16) & 0xffu);
                            if (PHONG) {
                                float gr[3];
                                if (tt == 0.0f) {
                                    clip_border_gradient(a, px, py, ss, dir, gr);
                                } else if (FILTER == 0) {
                                    corners(pp, a.dims, cc);
                                    gradient_planes<T, WIDE>(f, a, uu, cc, gr);
                                } else if (!gradient_interior<T, WIDE>(f, a, uu, cc, vv, gr)) {
                                    gradient_planes<T, WIDE>(f, a, uu, cc, gr);
                                }
                                float pos[3], col[3] = {cr, cg, cb};
#pragma unroll
                                for (int i = 0; i < 3; ++i)
                                    pos[i] = (((ss[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                                phong_fast(a, gr, pos, col);
                                cr = col[0];
                                cg = col[1];
                                cb = col[2];
                                ++n_shaded;
                            }
                            const float wgt = (1.0f - A) * ca;
                            r = r + wgt * cr;
                            g = g + wgt * cg;
                            bb = bb + wgt * cb;
                            A = A + (1.0f - A) * ca;
                        }
                        if (A >= kErtOpacity) {
                            A = 1.0f;
                            finished = true;
                        }
                    };
                    composite(v0, c0, s, u, vp, t);
                    if (pair && !finished) {
                        if (COUNT) {
                            ++n_slots;
                            n_shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
                            shade_now = false;
                        }
                        ++loop;
                        ++n_samples;
                        t = t1;
                        composite(v1, c1, s1, u1, vp1, t1);
                    }
                }
                if (COUNT) {
                    ++n_slots;
                    n_shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
                }
                t += a.step;
                finished = finished || (t > tend);
            }
            res = make_float4(r, g, bb, A);
        }
        a.out[(size_t)gy * a.grid_w + gx] = res;
    }
    if (COUNT) {
        wave_add(a.counters + 0, n_rays);
        wave_add(a.counters + 1, n_samples);
        wave_add(a.counters + 2, n_fetched);
        wave_add(a.counters + 3, n_shaded);
        wave_add_max64(a.counters + 4, n_slots);
        wave_add_max64(a.counters + 5, n_shade_slots);
    }
}

template <typename T, bool WIDE, int FILTER, bool PHONG>
static void launch_t(const MarchArgs& a, bool count, hipStream_t st) {
    const uint32_t tx = (a.grid_w + 15) / 16, ty = (a.grid_h + 15) / 16;
    const dim3 grid(xcd_tile_grid(tx, ty));
    if (count)
        hipLaunchKernelGGL((march_fast_kernel<T, WIDE, FILTER, PHONG, true>), grid, dim3(256), 0, st,
                           a, tx, ty);
    else
        hipLaunchKernelGGL((march_fast_kernel<T, WIDE, FILTER, PHONG, false>), grid, dim3(256), 0, st,
                           a, tx, ty);
}

template <typename T, bool WIDE>
static void launch_w(const MarchArgs& a, int filter, bool phong, bool count, hipStream_t st) {
    if (filter == 0) {
        if (phong) launch_t<T, WIDE, 0, true>(a, count, st);
        else launch_t<T, WIDE, 0, false>(a, count, st);
    } else {
        if (phong) launch_t<T, WIDE, 1, true>(a, count, st);
        else launch_t<T, WIDE, 1, false>(a, count, st);
    }
}

void launch_march_fast(const MarchArgs& a, int dtype, int filter, bool phong, bool count,
                       bool wide, hipStream_t st) {
    if (dtype == 0) {
        if (wide) launch_w<uint8_t, true>(a, filter, phong, count, st);
        else launch_w<uint8_t, false>(a, filter, phong, count, st);
    } else {
        if (wide) launch_w<uint16_t, true>(a, filter, phong, count, st);
        else launch_w<uint16_t, false>(a, filter, phong, count, st);
    }
}

}  // namespace hipdvr
