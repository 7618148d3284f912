// march_cell.h -- device code of the cell-record ray march, shared by the
// per-frame kernel (march_cell.hip) and the batched persistent kernel
// (batch.hip).
//
// The record of cell (x, y, z) holds the eight voxels of its trilinear
// footprint, (x..x+1, y..y+1, z..z+1) clamped to the volume, in the order the
// trilinear sum visits them (llB, lrB, urB, ulB, llF, lrF, urF, ulF of
// VolumeRAM::getVoxelNormalizedLinear, volumeram.cpp:78-97).  One sample is
// one 16-byte (uint16) or 8-byte (uint8) load instead of eight 2- or 1-byte
// gathers: the march is bound by the vector-memory issue rate per load
// instruction, not by bytes, so one wide load per sample is the MI355X trade
// of HBM capacity (8x the voxel bytes) for load instructions.
//
// Sample positions, corners, weights, the 8-term sum, TF classification,
// opacity correction, Phong's six central-difference lookups
// (mod_gradients.frag:202-214) and the lighting are the same expressions as
// the reference-layout kernel (march_ref.hip, restating
// CPURaycaster::directRendering, cpuraycaster.cpp:210-326), so the image is
// bit-identical to it, shaded or not.  Empty-space skipping is the exact
// brick fast-forward of the bricked kernel; the samples after a non-empty
// sample are fetched K at a time (their loads in flight together) and
// composited in order; a sample fetched past early termination is dropped.
#pragma once

#include <type_traits>

#include "cell_layout.h"
#include "march_common.h"
#include "step_run.h"

namespace hipdvr {

template <typename T>
struct CellRecord;
template <>
struct CellRecord<uint16_t> {
    typedef uint4 type;
    // stored values as floats (no normalization)
    static __device__ __forceinline__ void raw(const uint4& r, float v[8]) {
        v[0] = (float)(r.x & 0xffffu);
        v[1] = (float)(r.x >> 16);
        v[2] = (float)(r.y & 0xffffu);
        v[3] = (float)(r.y >> 16);
        v[4] = (float)(r.z & 0xffffu);
        v[5] = (float)(r.z >> 16);
        v[6] = (float)(r.w & 0xffffu);
        v[7] = (float)(r.w >> 16);
    }
    static __device__ __forceinline__ void unpack(const uint4& r, float v[8]) {
        v[0] = VoxNorm<uint16_t>::apply(r.x & 0xffffu);
        v[1] = VoxNorm<uint16_t>::apply(r.x >> 16);
        v[2] = VoxNorm<uint16_t>::apply(r.y & 0xffffu);
        v[3] = VoxNorm<uint16_t>::apply(r.y >> 16);
        v[4] = VoxNorm<uint16_t>::apply(r.z & 0xffffu);
        v[5] = VoxNorm<uint16_t>::apply(r.z >> 16);
        v[6] = VoxNorm<uint16_t>::apply(r.w & 0xffffu);
        v[7] = VoxNorm<uint16_t>::apply(r.w >> 16);
    }
};
template <>
struct CellRecord<uint8_t> {
    typedef uint2 type;
    // stored values as floats (no normalization)
    static __device__ __forceinline__ void raw(const uint2& r, float v[8]) {
        v[0] = (float)(r.x & 0xffu);
        v[1] = (float)((r.x >> 8) & 0xffu);
        v[2] = (float)((r.x >> 16) & 0xffu);
        v[3] = (float)(r.x >> 24);
        v[4] = (float)(r.y & 0xffu);
        v[5] = (float)((r.y >> 8) & 0xffu);
        v[6] = (float)((r.y >> 16) & 0xffu);
        v[7] = (float)(r.y >> 24);
    }
    static __device__ __forceinline__ void unpack(const uint2& r, float v[8]) {
        v[0] = VoxNorm<uint8_t>::apply(r.x & 0xffu);
        v[1] = VoxNorm<uint8_t>::apply((r.x >> 8) & 0xffu);
        v[2] = VoxNorm<uint8_t>::apply((r.x >> 16) & 0xffu);
        v[3] = VoxNorm<uint8_t>::apply(r.x >> 24);
        v[4] = VoxNorm<uint8_t>::apply(r.y & 0xffu);
        v[5] = VoxNorm<uint8_t>::apply((r.y >> 8) & 0xffu);
        v[6] = VoxNorm<uint8_t>::apply((r.y >> 16) & 0xffu);
        v[7] = VoxNorm<uint8_t>::apply(r.y >> 24);
    }
};

// Record addressing (cell_layout.h): 2x2x2 cells per group (one 128-byte
// line for uint16), groups x-fastest inside bricks of kCellBrick^3 groups,
// bricks x-fastest.
template <typename T, bool WIDE>
struct CellFetch {
    typedef typename CellRecord<T>::type rec_t;
    typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type off_t;
    const rec_t* p;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t nbx, nby;   // bricks along x and y

    __device__ __forceinline__ off_t index(const uint32_t lo[3]) const { return cell_index<off_t>(lo, nbx, nby); }
    // brick stride of axis i, in records
    __device__ __forceinline__ off_t brick_stride(int i) const {
        const off_t one = (off_t)1 << kCellBrickRecLog2;
        return i == 0 ? one : (i == 1 ? one * nbx : one * nbx * nby);
    }
    __device__ __forceinline__ rec_t load(off_t i) const {
        if constexpr (WIDE) {
            return p[i];
        } else if constexpr (sizeof(T) == 2) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (uint32_t)i * 16u, 0, 0);
            return rec_t{v[0], v[1], v[2], v[3]};
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, (uint32_t)i * 8u, 0, 0);
            rec_t r;
            r.x = v[0];
            r.y = v[1];
            return r;
        }
    }
};

// Voxel position, corners and record index of the texture-space point s.
template <typename T, bool WIDE>
__device__ __forceinline__ typename CellFetch<T, WIDE>::off_t cell_of(const CellFetch<T, WIDE>& f,
                                                                      const MarchArgs& a, const float s[3],
                                                                      float p[3]) {
    // corners() of volumeram.cpp:79-86 on a clamped position (vp >= 0, so
    // max(vp, 0) == vp): the record holds the clamped upper corners, so only
    // the lower corner and the fraction are needed
    uint32_t lo[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float vp = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
        p[i] = vp - floorf(vp);
        const uint32_t l = (uint32_t)vp, m = a.dims[i] - 1;
        lo[i] = l < m ? l : m;
    }
    return f.index(lo);
}

// Central differences of six trilinear lookups at s +- one voxel per axis,
// as march_ref.hip computes them; the six loads are issued together.  RAW
// (kernels whose shading is checked by tolerance): the lookups interpolate
// the stored values and the difference is scaled once by grad_raw, equal up
// to float rounding.
template <typename T, bool WIDE, bool RAW = false>
__device__ __forceinline__ void cell_gradient(const CellFetch<T, WIDE>& f, const MarchArgs& a,
                                              const float s[3], float gr[3]) {
    typename CellFetch<T, WIDE>::rec_t rec[6];
    float p[6][3];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int axis = k % 3;
        const float off = (k < 3) ? a.grad_off[axis] : -a.grad_off[axis];
        float q[3] = {s[0], s[1], s[2]};
        q[axis] = s[axis] + off;
        q[(axis + 1) % 3] = s[(axis + 1) % 3] + 0.0f;
        q[(axis + 2) % 3] = s[(axis + 2) % 3] + 0.0f;
        rec[k] = f.load(cell_of<T, WIDE>(f, a, q, p[k]));
    }
    float vals[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float v[8];
        if (RAW) CellRecord<T>::raw(rec[k], v);
        else CellRecord<T>::unpack(rec[k], v);
        vals[k] = trilerp(v, p[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (RAW) gr[i] = (vals[3 + i] - vals[i]) * a.grad_raw[i];
        else gr[i] = ((vals[3 + i] - vals[i]) * a.grad_scale[i]) * a.rwm_scale;
    }
}

// Interior central differences from the cell records: when every axis has
// u_i in [1, dim_i - 3) and a proper fraction, the six lookups at +-1 voxel
// have the sample's own weights p, and their footprints are the sample's
// corner planes plus the outer planes lo-1 and lo+2 (records of the cells
// lo -+ e_i).  Equal to the six-lookup form up to float rounding (for the
// kernels whose shading is checked by tolerance): the differences are taken
// in stored units and scaled once by grad_raw = grad_scale * rwm_scale /
// range, the four-term sums use fused multiply-adds, and the neighbour
// records are found from the sample's own record index `own` (its low three
// bits are the parities of lo; a step to the neighbouring group moves by the
// group stride inside a brick, else to the neighbouring brick's facing group;
// cell_layout.h).  Returns false (and leaves gr) outside the interior.
template <typename T, bool WIDE>
__device__ __forceinline__ bool cell_gradient_interior(const CellFetch<T, WIDE>& f, const MarchArgs& a,
                                                       const float s[3],
                                                       const typename CellFetch<T, WIDE>::rec_t& rec,
                                                       typename CellFetch<T, WIDE>::off_t own, const float p[3],
                                                       float gr[3]) {
    typedef typename CellFetch<T, WIDE>::off_t off_t;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float u = (a.fdims[i] * s[i]) + (-0.5f);
        if (!(u >= 1.0f) || !(u + 3.0f < a.fdims[i]) || !(p[i] > 0.0f)) return false;
    }
    off_t nb[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) cell_neighbours<off_t>(own, i, f.brick_stride(i), nb[i], nb[3 + i]);
    typename CellFetch<T, WIDE>::rec_t nr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) nr[k] = f.load(nb[k]);
    float v[8], m[3][8], q[3][8];
    CellRecord<T>::raw(rec, v);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        CellRecord<T>::raw(nr[i], m[i]);
        CellRecord<T>::raw(nr[3 + i], q[i]);
    }
    const float px = p[0], py = p[1], pz = p[2];
    const float qx = 1.f - px, qy = 1.f - py, qz = 1.f - pz;
    auto dot4 = [](float a0, float a1, float a2, float a3, const float w[4]) {
        return __builtin_fmaf(a3, w[3], __builtin_fmaf(a2, w[2], __builtin_fmaf(a1, w[1], a0 * w[0])));
    };
    // (qa*bm + pa*b0) - (qa*b1 + pa*bp): the lookup at -1 voxel minus the one at +1
    auto diff = [](float qa, float pa, float bm, float b0, float b1, float bp) {
        return __builtin_fmaf(-pa, bp, __builtin_fmaf(-qa, b1, __builtin_fmaf(pa, b0, qa * bm)));
    };
    // corner order: 0 x0y0z0, 1 x1y0z0, 2 x1y1z0, 3 x0y1z0, 4 x0y0z1, 5 x1y0z1, 6 x1y1z1, 7 x0y1z1
    {
        const float w[4] = {qy * qz, py * qz, qy * pz, py * pz};
        const float bm = dot4(m[0][0], m[0][3], m[0][4], m[0][7], w);
        const float b0 = dot4(v[0], v[3], v[4], v[7], w);
        const float b1 = dot4(v[1], v[2], v[5], v[6], w);
        const float bp = dot4(q[0][1], q[0][2], q[0][5], q[0][6], w);
        gr[0] = diff(qx, px, bm, b0, b1, bp) * a.grad_raw[0];
    }
    {
        const float w[4] = {qx * qz, px * qz, qx * pz, px * pz};
        const float bm = dot4(m[1][0], m[1][1], m[1][4], m[1][5], w);
        const float b0 = dot4(v[0], v[1], v[4], v[5], w);
        const float b1 = dot4(v[3], v[2], v[7], v[6], w);
        const float bp = dot4(q[1][3], q[1][2], q[1][7], q[1][6], w);
        gr[1] = diff(qy, py, bm, b0, b1, bp) * a.grad_raw[1];
    }
    {
        const float w[4] = {qx * qy, px * qy, qx * py, px * py};
        const float bm = dot4(m[2][0], m[2][1], m[2][3], m[2][2], w);
        const float b0 = dot4(v[0], v[1], v[3], v[2], w);
        const float b1 = dot4(v[4], v[5], v[7], v[6], w);
        const float bp = dot4(q[2][4], q[2][5], q[2][7], q[2][6], w);
        gr[2] = diff(qz, pz, bm, b0, b1, bp) * a.grad_raw[2];
    }
    return true;
}

// Per-ray counters of a march (COUNT builds).
struct RayCounters {
    unsigned long long rays = 0, samples = 0, fetched = 0, shaded = 0, slots = 0, shade_slots = 0;
};

// Occupancy bit words staged in LDS by the per-frame kernel up to this count
// (4 KiB: 32768 bricks, e.g. 512^3 at 16^3 bricks).
constexpr uint32_t kMaxLdsOccWords = 1024;

// One ray of the cell march from entry point e to exit point l (neither the
// background): returns the composited RGBA.  SHADE: 0 none, 1 Phong with the
// interior gradient and the hardware-approximate lighting terms (within the
// oracle's tolerance), 2 Phong bit-identical to the reference-layout kernel.  s_tf / s_alpha are the TF and
// the opacity-correction table in LDS; a may live in kernel arguments or in
// global memory (batched launches).
// The sample loop of one ray from state (t, loop, r, g, bb, A) up to sample
// index cap (exclusive) or the ray's end; returns true when the ray is
// finished (past tend, or terminated early).
template <typename T, bool WIDE, int SHADE, bool COUNT, int K>
__device__ __forceinline__ bool cell_march(const MarchArgs& a, const CellFetch<T, WIDE>& f, const uint32_t* s_tf,
                                           const float* s_alpha, const uint32_t* s_occ, int px, int py,
                                           const float first[3], const float dir[3], float tend, int cap, float& t,
                                           int& loop, float& r, float& g, float& bb, float& A, RayCounters& n) {
    typedef typename CellFetch<T, WIDE>::rec_t rec_t;
    const uint32_t bl = a.brick_log2;
    const uint32_t nbx = a.nbricks[0], nbxy = a.nbricks[0] * a.nbricks[1];
    uint32_t cur_brick = 0xffffffffu;
    bool cur_occ = false;
    float t_lim = 0.0f;
    // reciprocals for the empty-brick bound (a bound with a 0.05-voxel
    // margin needs no correctly rounded quotient); an axis with |dir| below
    // 1e-30 cannot move 0.05 voxel within a ray and bounds nothing
    float rdir[3], rfd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rdir[i] = fabsf(dir[i]) < 1e-30f ? 0.0f : __builtin_amdgcn_rcpf(dir[i]);
        rfd[i] = __builtin_amdgcn_rcpf(a.fdims[i]);
    }
    bool finished = false;
    while (!finished && loop < cap) {
        const float s[3] = {first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2]};
        float vp[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
        const uint32_t brick = (((uint32_t)vp[2] >> bl) * nbxy) + (((uint32_t)vp[1] >> bl) * nbx) +
                               ((uint32_t)vp[0] >> bl);
        if (brick != cur_brick) {
            cur_brick = brick;
            // occupancy bits staged in LDS when they fit, else the byte table
            cur_occ = s_occ ? ((s_occ[brick >> 5] >> (brick & 31)) & 1u) != 0 : a.brick_occ[brick] != 0;
            if (!cur_occ) t_lim = empty_brick_limit_rcp(a, first, rdir, rfd, vp, bl);
        }
        if (!cur_occ) {
            // every sample at t' < t_lim lies in this empty brick: only
            // the count and the float sequence t += step remain
            ++n.samples;
            if (COUNT) ++n.slots;
            {
                // most of the run in closed form; the last steps below take
                // the reference's float sequence step by step
                float d = 0.0f;
                const int j = closed_form_steps(t, a.step, fminf(t_lim, tend), cap - loop, d);
                if (j > 0) {
                    t = t + (float)j * d;
                    loop += j;
                    n.samples += (unsigned long long)j;
                    if (COUNT) n.slots += (unsigned long long)j;
                }
            }
            float tn = t + a.step;
            while (tn < t_lim && !(tn > tend) && loop + 1 < cap) {
                t = tn;
                ++loop;
                ++n.samples;
                if (COUNT) ++n.slots;
                tn = t + a.step;
            }
            ++loop;
            t = tn;
            finished = t > tend;
            continue;
        }
        // K samples t, t+step, ... (each within the ray and the cap);
        // all loads issued before the first is composited
        float tk[K], pk[K][3];
        rec_t rk[K];
        typename CellFetch<T, WIDE>::off_t ik[K];
        bool valid[K];
        {
            float tt = t;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                valid[k] = (k == 0) || (valid[k - 1] && !(tt > tend) && loop + k < cap);
                tk[k] = tt;
                if (valid[k]) {
                    float sk[3];
                    if (k == 0) {
                        sk[0] = s[0];
                        sk[1] = s[1];
                        sk[2] = s[2];
                    } else {
                        sk[0] = first[0] + tt * dir[0];
                        sk[1] = first[1] + tt * dir[1];
                        sk[2] = first[2] + tt * dir[2];
                    }
                    ik[k] = cell_of<T, WIDE>(f, a, sk, pk[k]);
                    rk[k] = f.load(ik[k]);
                }
                tt = tt + a.step;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!valid[k] || finished) break;
            bool shade_now = false;
            ++loop;
            ++n.samples;
            ++n.fetched;
            float v[8];
            CellRecord<T>::unpack(rk[k], v);
            const float inten = trilerp(v, pk[k]);
            const uint32_t c8 = s_tf[tf_index(inten, a)];
            const float ca = s_alpha[c8 >> 24];
            if (ca > 0.0f) {
                float col[3] = {byte_norm(c8 & 0xffu), byte_norm((c8 >> 8) & 0xffu),
                                byte_norm((c8 >> 16) & 0xffu)};
                if (SHADE) {
                    shade_now = true;
                    const float tt = tk[k];
                    const float ss[3] = {first[0] + tt * dir[0], first[1] + tt * dir[1],
                                         first[2] + tt * dir[2]};
                    float gr[3];
                    if (tt == 0.0f) clip_border_gradient(a, px, py, ss, dir, gr);
                    else if (SHADE == 2) cell_gradient<T, WIDE>(f, a, ss, gr);
                    else if (!cell_gradient_interior<T, WIDE>(f, a, ss, rk[k], ik[k], pk[k], gr))
                        cell_gradient<T, WIDE, true>(f, a, ss, gr);
                    float pos[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        pos[i] = (((ss[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                    if (SHADE == 2) shade_model(a, gr, pos, col);
                    else phong_fast(a, gr, pos, col);
                    ++n.shaded;
                }
                const float wgt = (1.0f - A) * ca;
                r = r + wgt * col[0];
                g = g + wgt * col[1];
                bb = bb + wgt * col[2];
                A = A + (1.0f - A) * ca;
            }
            if (COUNT) {
                ++n.slots;
                n.shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
            }
            if (A >= kErtOpacity) {
                A = 1.0f;
                finished = true;
            }
            t = tk[k] + a.step;
            finished = finished || (t > tend);
        }
    }
    return finished;
}

template <typename T, bool WIDE, int SHADE, bool COUNT, int K>
__device__ __forceinline__ float4 cell_ray(const MarchArgs& a, const CellFetch<T, WIDE>& f, const uint32_t* s_tf,
                                           const float* s_alpha, const uint32_t* s_occ, int px, int py, float4 e,
                                           float4 l, RayCounters& n) {
    ++n.rays;
    const float first[3] = {e.x, e.y, e.z};
    const float last[3] = {l.x, l.y, l.z};
    float dir[3];
    const float tend = ray_setup(first, last, a.step, dir);
    float t = 0.0f, r = 0.f, g = 0.f, bb = 0.f, A = 0.f;
    int loop = 0;   // samples taken (the reference loop counter)
    cell_march<T, WIDE, SHADE, COUNT, K>(a, f, s_tf, s_alpha, s_occ, px, py, first, dir, tend, a.loop_cap, t, loop, r,
                                         g, bb, A, n);
    return make_float4(r, g, bb, A);
}

// Cell-record addressing of the volume in a (buffer descriptor when narrow).
template <typename T, bool WIDE>
__device__ __forceinline__ CellFetch<T, WIDE> cell_fetch_of(const MarchArgs& a) {
    typedef typename CellFetch<T, WIDE>::rec_t rec_t;
    CellFetch<T, WIDE> f;
    f.p = (const rec_t*)a.cells;
    f.nbx = cell_bricks(a.dims[0]);
    f.nby = cell_bricks(a.dims[1]);
    if (!WIDE) {
        const uint32_t dims[3] = {a.dims[0], a.dims[1], a.dims[2]};
        const uint32_t bytes = (uint32_t)(cell_record_count(dims) * sizeof(rec_t));
        f.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.cells, (short)0, (int)bytes, 0x00020000);
    }
    return f;
}

}  // namespace hipdvr
