// march_cell.hip -- per-frame launches of the cell-record ray march (the
// device code and its parity argument are in march_cell.h): one workgroup per
// 16x16 pixel tile in XCD-aware order, and a software-pipelined variant.
#include "kernels.h"
#include "march_cell.h"
#include "march_common.h"

namespace hipdvr {

template <typename T, bool WIDE, bool PHONG, bool COUNT, int K>
__global__ __launch_bounds__(256) void march_cell_kernel(const MarchArgs a, const uint32_t tiles_x,
                                                         const uint32_t tiles_total,
                                                         const uint32_t tiles_per_xcd) {
    // XCD-aware tile order: workgroup b runs on XCD b % 8, which marches the
    // contiguous band of tiles [k*tiles_per_xcd, (k+1)*tiles_per_xcd).
    const uint32_t b = blockIdx.x;
    const uint32_t tile = (b & 7u) * tiles_per_xcd + (b >> 3);
    if (tile >= tiles_total) return;   // whole workgroup leaves together
    if (a.cond_flag && *a.cond_flag == 0) return;
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int gx = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int gy = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int px = gx * a.stride + a.pix_offset, py = gy * a.stride + a.pix_offset;
    const bool inside = (gx < a.grid_w) && (gy < a.grid_h) && (px < a.w) && (py < a.h);
    const CellFetch<T, WIDE> f = cell_fetch_of<T, WIDE>(a);
    RayCounters n;
    if (inside) {
        const size_t pix = (size_t)py * a.w + px;
        const float4 e = a.entry[pix];
        const float4 l = a.exit[pix];
        float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!is_background(e, l)) res = cell_ray<T, WIDE, PHONG, COUNT, K>(a, f, s_tf, s_alpha, px, py, e, l, n);
        a.out[(size_t)gy * a.grid_w + gx] = res;
    }
    if (COUNT) {
        wave_add(a.counters + 0, n.rays);
        wave_add(a.counters + 1, n.samples);
        wave_add(a.counters + 2, n.fetched);
        wave_add(a.counters + 3, n.shaded);
        wave_add_max64(a.counters + 4, n.slots);
        wave_add_max64(a.counters + 5, n.shade_slots);
    }
}

// Samples t_k = t_0 + k*step (the float sequence), k < K, fetched together:
// sample k is valid while the ray and the loop cap allow it.
template <typename rec_t, int K>
struct CellChunk {
    float t[K];
    float p[K][3];
    rec_t r[K];
    bool v[K];
};

template <typename T, bool WIDE, int K>
__device__ __forceinline__ uint32_t fetch_chunk(const CellFetch<T, WIDE>& f, const MarchArgs& a,
                                                const float first[3], const float dir[3], float t0, bool v0,
                                                int loop0, float tend, int cap,
                                                CellChunk<typename CellFetch<T, WIDE>::rec_t, K>& c) {
    float tt = t0;
    uint32_t brick0 = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        c.v[k] = (k == 0) ? v0 : (c.v[k - 1] && !(tt > tend) && loop0 + k < cap);
        c.t[k] = tt;
        if (c.v[k]) {
            const float sk[3] = {first[0] + tt * dir[0], first[1] + tt * dir[1], first[2] + tt * dir[2]};
            float vp[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(sk[i], a.fdims[i], a.fdm1[i]);
            if (k == 0) {
                const uint32_t bl = a.brick_log2;
                brick0 = (((uint32_t)vp[2] >> bl) * (a.nbricks[0] * a.nbricks[1])) +
                         (((uint32_t)vp[1] >> bl) * a.nbricks[0]) + ((uint32_t)vp[0] >> bl);
            }
            Corners cc;
            corners(vp, a.dims, cc);
#pragma unroll
            for (int i = 0; i < 3; ++i) c.p[k][i] = cc.p[i];
            c.r[k] = f.load(f.index(cc.lo));
        }
        tt = tt + a.step;
    }
    return brick0;
}

// Software-pipelined variant: while chunk A is composited, the loads of the
// next chunk B are in flight; the brick occupancy bits sit in LDS when they
// fit (OCC_LDS).  Same samples, same order, same arithmetic as
// march_cell_kernel: bit-identical output.
template <typename T, bool WIDE, bool PHONG, bool COUNT, int K, bool OCC_LDS>
__global__ __launch_bounds__(256) void march_cell_pipe_kernel(const MarchArgs a, const uint32_t tiles_x,
                                                              const uint32_t tiles_total,
                                                              const uint32_t tiles_per_xcd) {
    const uint32_t b = blockIdx.x;
    const uint32_t tile = (b & 7u) * tiles_per_xcd + (b >> 3);
    if (tile >= tiles_total) return;
    if (a.cond_flag && *a.cond_flag == 0) return;
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    extern __shared__ uint32_t s_occ[];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    if (OCC_LDS)
        for (uint32_t i = tid; i < a.occ_words; i += 256) s_occ[i] = a.occ_bits[i];
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int gx = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int gy = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int px = gx * a.stride + a.pix_offset, py = gy * a.stride + a.pix_offset;
    const bool inside = (gx < a.grid_w) && (gy < a.grid_h) && (px < a.w) && (py < a.h);

    typedef typename CellFetch<T, WIDE>::rec_t rec_t;
    CellFetch<T, WIDE> f;
    f.p = (const rec_t*)a.cells;
    const uint32_t GX = (a.dims[0] + 1) >> 1, GY = (a.dims[1] + 1) >> 1, GZ = (a.dims[2] + 1) >> 1;
    f.sy = GX << 3;
    f.sz = f.sy * GY;
    if (!WIDE) {
        const uint32_t bytes = f.sz * GZ * (uint32_t)sizeof(rec_t);
        f.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.cells, (short)0, (int)bytes, 0x00020000);
    }
    const uint32_t bl = a.brick_log2;
    const uint32_t nbx = a.nbricks[0], nbxy = a.nbricks[0] * a.nbricks[1];
    auto occupied = [&](uint32_t brick) -> bool {
        if (OCC_LDS) return ((s_occ[brick >> 5] >> (brick & 31)) & 1u) != 0;
        return a.brick_occ[brick] != 0;
    };

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
            bool finished = false;
            int loop = 0;
            while (!finished && loop < cap) {
                const float s[3] = {first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2]};
                float vp[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
                const uint32_t brick = (((uint32_t)vp[2] >> bl) * nbxy) + (((uint32_t)vp[1] >> bl) * nbx) +
                                       ((uint32_t)vp[0] >> bl);
                if (!occupied(brick)) {
                    // exact fast-forward through the empty brick
                    const float t_lim = empty_brick_limit(a, first, dir, vp, bl);
                    ++n_samples;
                    if (COUNT) ++n_slots;
                    float tn = t + a.step;
                    while (tn < t_lim && !(tn > tend) && loop + 1 < cap) {
                        t = tn;
                        ++loop;
                        ++n_samples;
                        if (COUNT) ++n_slots;
                        tn = t + a.step;
                    }
                    ++loop;
                    t = tn;
                    finished = t > tend;
                    continue;
                }
                CellChunk<rec_t, K> ca;
                fetch_chunk<T, WIDE, K>(f, a, first, dir, t, true, loop, tend, cap, ca);
                while (true) {
                    const float tb = ca.t[K - 1] + a.step;
                    const bool more = ca.v[K - 1] && !(tb > tend) && loop + K < cap;
                    CellChunk<rec_t, K> cb;
                    uint32_t brick_b = 0;
                    if (more) brick_b = fetch_chunk<T, WIDE, K>(f, a, first, dir, tb, true, loop + K, tend, cap, cb);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if (!ca.v[k] || finished) break;
                        bool shade_now = false;
                        ++loop;
                        ++n_samples;
                        ++n_fetched;
                        float v[8];
                        CellRecord<T>::unpack(ca.r[k], v);
                        const float inten = trilerp(v, ca.p[k]);
                        const uint32_t c8 = s_tf[tf_index(inten, a)];
                        const float alpha = s_alpha[c8 >> 24];
                        if (alpha > 0.0f) {
                            float col[3] = {byte_norm(c8 & 0xffu), byte_norm((c8 >> 8) & 0xffu),
                                            byte_norm((c8 >> 16) & 0xffu)};
                            if (PHONG) {
                                shade_now = true;
                                const float tt = ca.t[k];
                                const float ss[3] = {first[0] + tt * dir[0], first[1] + tt * dir[1],
                                                     first[2] + tt * dir[2]};
                                float gr[3];
                                if (tt == 0.0f) clip_border_gradient(a, px, py, ss, dir, gr);
                                else cell_gradient<T, WIDE>(f, a, ss, gr);
                                float pos[3];
#pragma unroll
                                for (int i = 0; i < 3; ++i)
                                    pos[i] = (((ss[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                                phong(a, gr, pos, col);
                                ++n_shaded;
                            }
                            const float wgt = (1.0f - A) * alpha;
                            r = r + wgt * col[0];
                            g = g + wgt * col[1];
                            bb = bb + wgt * col[2];
                            A = A + (1.0f - A) * alpha;
                        }
                        if (COUNT) {
                            ++n_slots;
                            n_shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
                        }
                        if (A >= kErtOpacity) {
                            A = 1.0f;
                            finished = true;
                        }
                        t = ca.t[k] + a.step;
                        finished = finished || (t > tend);
                    }
                    if (finished || !more) break;
                    ca = cb;
                    if (!occupied(brick_b)) break;   // t == tb: the outer loop skips the empty brick
                }
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

template <typename T, bool WIDE, bool PHONG, int K, bool OCC_LDS>
static void launch_pk(const MarchArgs& a, bool count, hipStream_t st) {
    const uint32_t tx = (a.grid_w + 15) / 16, ty = (a.grid_h + 15) / 16;
    const uint32_t total = tx * ty;
    const uint32_t per_xcd = (total + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const size_t shm = OCC_LDS ? (size_t)a.occ_words * 4 : 0;
    if (count)
        hipLaunchKernelGGL((march_cell_pipe_kernel<T, WIDE, PHONG, true, K, OCC_LDS>), grid, dim3(256), shm, st, a,
                           tx, total, per_xcd);
    else
        hipLaunchKernelGGL((march_cell_pipe_kernel<T, WIDE, PHONG, false, K, OCC_LDS>), grid, dim3(256), shm, st,
                           a, tx, total, per_xcd);
}

template <typename T, bool WIDE, bool PHONG>
static void launch_pb(const MarchArgs& a, bool count, int batch, hipStream_t st) {
    const bool lds = a.occ_words <= kMaxLdsOccWords;
    if (batch <= 2) {
        if (lds) launch_pk<T, WIDE, PHONG, 2, true>(a, count, st);
        else launch_pk<T, WIDE, PHONG, 2, false>(a, count, st);
    } else {
        if (lds) launch_pk<T, WIDE, PHONG, 4, true>(a, count, st);
        else launch_pk<T, WIDE, PHONG, 4, false>(a, count, st);
    }
}

void launch_march_cell_pipe(const MarchArgs& a, int dtype, bool phong, bool count, bool wide, int batch,
                            hipStream_t st) {
    if (dtype == 0) {
        if (wide) {
            if (phong) launch_pb<uint8_t, true, true>(a, count, batch, st);
            else launch_pb<uint8_t, true, false>(a, count, batch, st);
        } else {
            if (phong) launch_pb<uint8_t, false, true>(a, count, batch, st);
            else launch_pb<uint8_t, false, false>(a, count, batch, st);
        }
    } else {
        if (wide) {
            if (phong) launch_pb<uint16_t, true, true>(a, count, batch, st);
            else launch_pb<uint16_t, true, false>(a, count, batch, st);
        } else {
            if (phong) launch_pb<uint16_t, false, true>(a, count, batch, st);
            else launch_pb<uint16_t, false, false>(a, count, batch, st);
        }
    }
}

template <typename T, bool WIDE, bool PHONG, int K>
static void launch_k(const MarchArgs& a, bool count, hipStream_t st) {
    const uint32_t tx = (a.grid_w + 15) / 16, ty = (a.grid_h + 15) / 16;
    const uint32_t total = tx * ty;
    const uint32_t per_xcd = (total + 7) / 8;
    const dim3 grid(per_xcd * 8);
    if (count)
        hipLaunchKernelGGL((march_cell_kernel<T, WIDE, PHONG, true, K>), grid, dim3(256), 0, st, a, tx, total,
                           per_xcd);
    else
        hipLaunchKernelGGL((march_cell_kernel<T, WIDE, PHONG, false, K>), grid, dim3(256), 0, st, a, tx, total,
                           per_xcd);
}

template <typename T, bool WIDE>
static void launch_cw(const MarchArgs& a, bool phong, bool count, int batch, hipStream_t st) {
    if (phong) {
        if (batch <= 2) launch_k<T, WIDE, true, 2>(a, count, st);
        else if (batch <= 4) launch_k<T, WIDE, true, 4>(a, count, st);
        else launch_k<T, WIDE, true, 8>(a, count, st);
    } else {
        if (batch <= 2) launch_k<T, WIDE, false, 2>(a, count, st);
        else if (batch <= 4) launch_k<T, WIDE, false, 4>(a, count, st);
        else launch_k<T, WIDE, false, 8>(a, count, st);
    }
}

void launch_march_cell(const MarchArgs& a, int dtype, bool phong, bool count, bool wide, int batch,
                       hipStream_t st) {
    if (dtype == 0) {
        if (wide) launch_cw<uint8_t, true>(a, phong, count, batch, st);
        else launch_cw<uint8_t, false>(a, phong, count, batch, st);
    } else {
        if (wide) launch_cw<uint16_t, true>(a, phong, count, batch, st);
        else launch_cw<uint16_t, false>(a, phong, count, batch, st);
    }
}

}  // namespace hipdvr
