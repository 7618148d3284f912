// march_slab.h -- the slab-staged ray march: each workgroup (one 16x16 pixel
// tile) stages the part of the volume its rays need next into LDS, one slab
// of planes at a time, and reconstructs every sample from LDS.
//
// Why: the rays of a tile are nearly parallel (a 16-pixel tile of a 1024^2
// view spans half a degree), so between two planes across their common main
// axis they all lie in one small box of voxels -- about 20 x 20 x 11 for the
// 512^3 / 1024^2 workload.  Staged once per slab with coalesced row loads of
// the linear (VolumeRAM, x-fastest) volume, the box serves every trilinear
// sample and every Phong gradient of the tile's rays in that slab from LDS:
// no replicated records in HBM (the volume is read as it is stored), no
// per-sample gathers through L2.
//
// Scheduling.  A ray needs data from slab s when its next sample lies in a
// brick that may map to a non-zero TF alpha and the sample's lower corner
// along the slab axis d lies in planes [s*S, s*S + S).  Samples in empty
// bricks need no data: they are fast-forwarded exactly as in the other
// kernels (closed form plus the reference's t += step) whatever the slab.
// Each round the workgroup takes the nearest slab any ray needs (in the
// tiles's direction along d), stages the box bounding those rays' segments
// in it (plus one plane before and two after: the gradient's neighbours),
// and those rays march until their next data sample lies outside the slab.
// A ray's corner along d is monotonic in t, so every ray meets its slabs in
// order and the loop ends.  A sample whose footprint is outside the staged
// box (a box clipped to the LDS capacity, a ray running against the tile's
// direction) reads the voxels from global memory instead: the box decides
// speed, never the result.
//
// Arithmetic: positions, corners, weights, the 8-term sum, the TF lookup,
// opacity correction, compositing, the skip logic and the Phong gradient
// forms are those of the cell kernel (march_cell.h) -- the same voxel values
// in the same expressions -- so the unshaded image equals the oracle bit for
// bit and the shaded one equals the cell kernel's.
#pragma once

#include <type_traits>

#include "march_cell.h"
#include "march_common.h"
#include "step_run.h"

namespace hipdvr {

constexpr int kSlabLog2 = 3;                  // S = 8 planes per slab
constexpr int kSlab = 1 << kSlabLog2;
// LDS voxels of the staged box: 24 KiB for either voxel type (12288 uint16
// or 24576 uint8), beside the 16 KiB TF table, so three workgroups fit a CU.
template <typename T>
constexpr int kBoxCap = 24576 / (int)sizeof(T);
constexpr int kNoSlab = 0x7fffffff;           // key of a ray that needs no more data
constexpr int kAnySlab = -2;                  // slab argument: march to the end, whatever the slab
// Rounds after which every remaining ray marches to its end in one round
// with the box disabled (global reads): a bound on the loop whatever the
// rays' directions.
constexpr int kMaxSlabRounds = 1024;

// Staged box: voxels at virtual coordinates b0 + [0, e) per axis, x-fastest,
// each holding the voxel at the coordinate clamped to the volume (so the
// clamped upper corner of a footprint at the volume's last voxel is read
// like any other).  e[0] == 0: nothing staged.
struct SlabBox {
    int b0[3];
    int e[3];
    int sy, sz;        // e[0], e[0] * e[1]
};

// Voxel reads around lower corner lo: from the box when the footprint
// [lo + lo_off, lo + hi_off] fits in it on every axis, else from global
// memory with coordinates clamped to the volume.
template <typename T>
struct SlabReader {
    const T* box;          // LDS
    const T* vol;          // linear volume, global
    SlabBox g;
    uint32_t dims[3];

    __device__ __forceinline__ bool fits(const uint32_t lo[3], int lo_off, int hi_off) const {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int a = (int)lo[i] + lo_off - g.b0[i], b = (int)lo[i] + hi_off - g.b0[i];
            if (a < 0 || b >= g.e[i]) return false;
        }
        return true;
    }
    __device__ __forceinline__ int box_index(const uint32_t lo[3]) const {
        return ((int)lo[0] - g.b0[0]) + ((int)lo[1] - g.b0[1]) * g.sy + ((int)lo[2] - g.b0[2]) * g.sz;
    }
    __device__ __forceinline__ uint32_t gvox(int x, int y, int z) const {
        const int mx = (int)dims[0] - 1, my = (int)dims[1] - 1, mz = (int)dims[2] - 1;
        x = x < 0 ? 0 : (x > mx ? mx : x);
        y = y < 0 ? 0 : (y > my ? my : y);
        z = z < 0 ? 0 : (z > mz ? mz : z);
        return vol[((size_t)z * dims[1] + (size_t)y) * dims[0] + (size_t)x];
    }
    // The 8 voxels of the footprint (lo, lo + 1 clamped), in the order of
    // the reference's sum (llB, lrB, urB, ulB, llF, lrF, urF, ulF).
    __device__ __forceinline__ void corners8(const uint32_t lo[3], uint32_t v[8]) const {
        if (fits(lo, 0, 1)) {
            const int i = box_index(lo);
            v[0] = box[i];
            v[1] = box[i + 1];
            v[2] = box[i + g.sy + 1];
            v[3] = box[i + g.sy];
            v[4] = box[i + g.sz];
            v[5] = box[i + g.sz + 1];
            v[6] = box[i + g.sz + g.sy + 1];
            v[7] = box[i + g.sz + g.sy];
        } else {
            const int x = (int)lo[0], y = (int)lo[1], z = (int)lo[2];
            v[0] = gvox(x, y, z);
            v[1] = gvox(x + 1, y, z);
            v[2] = gvox(x + 1, y + 1, z);
            v[3] = gvox(x, y + 1, z);
            v[4] = gvox(x, y, z + 1);
            v[5] = gvox(x + 1, y, z + 1);
            v[6] = gvox(x + 1, y + 1, z + 1);
            v[7] = gvox(x, y + 1, z + 1);
        }
    }
};

// Corners (lower corner clamped to dim - 1, as the cell records hold them)
// and weights of the texture-space point s.
__device__ __forceinline__ void slab_corner(const MarchArgs& a, const float s[3], uint32_t lo[3], float p[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float vp = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
        p[i] = vp - floorf(vp);
        const uint32_t l = (uint32_t)vp, m = a.dims[i] - 1;
        lo[i] = l < m ? l : m;
    }
}

template <typename T>
__device__ __forceinline__ float slab_lookup(const SlabReader<T>& rd, const MarchArgs& a, const float s[3],
                                             bool raw) {
    uint32_t lo[3], v8[8];
    float p[3], v[8];
    slab_corner(a, s, lo, p);
    rd.corners8(lo, v8);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = raw ? (float)v8[k] : VoxNorm<T>::apply(v8[k]);
    return trilerp(v, p);
}

// Central differences of six trilinear lookups at s +- one voxel per axis
// (cell_gradient's forms: RAW differences of stored values scaled once, or
// the reference kernel's normalized ones).
template <typename T, bool RAW>
__device__ __forceinline__ void slab_gradient6(const SlabReader<T>& rd, const MarchArgs& a, const float s[3],
                                               float gr[3]) {
    float vals[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int axis = k % 3;
        const float off = (k < 3) ? a.grad_off[axis] : -a.grad_off[axis];
        float q[3] = {s[0], s[1], s[2]};
        q[axis] = s[axis] + off;
        q[(axis + 1) % 3] = s[(axis + 1) % 3] + 0.0f;
        q[(axis + 2) % 3] = s[(axis + 2) % 3] + 0.0f;
        vals[k] = slab_lookup<T>(rd, a, q, RAW);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (RAW) gr[i] = (vals[3 + i] - vals[i]) * a.grad_raw[i];
        else gr[i] = ((vals[3 + i] - vals[i]) * a.grad_scale[i]) * a.rwm_scale;
    }
}

// Interior central differences (cell_gradient_interior's form): the six
// lookups at +-1 voxel share the sample's weights p, their footprints are the
// sample's corner planes and the outer planes lo - 1 and lo + 2.  v: the
// sample's own 8 stored values.  Returns false outside the interior.
template <typename T>
__device__ __forceinline__ bool slab_gradient_interior(const SlabReader<T>& rd, const MarchArgs& a,
                                                       const float s[3], const uint32_t lo[3], const float p[3],
                                                       const uint32_t v8[8], float gr[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float u = (a.fdims[i] * s[i]) + (-0.5f);
        if (!(u >= 1.0f) || !(u + 3.0f < a.fdims[i]) || !(p[i] > 0.0f)) return false;
    }
    // outer planes: m[axis][4] at lo - 1, q[axis][4] at lo + 2 along the
    // axis, over the sample's 2x2 footprint of the other two axes, in the
    // corner order the dot products below use
    uint32_t m[3][4], q[3][4];
    const int x = (int)lo[0], y = (int)lo[1], z = (int)lo[2];
    if (rd.fits(lo, -1, 2)) {
        const int i = rd.box_index(lo), sy = rd.g.sy, sz = rd.g.sz;
        const T* b = rd.box;
        // x: (y, z) = (0,0), (1,0), (0,1), (1,1)
        m[0][0] = b[i - 1]; m[0][1] = b[i - 1 + sy]; m[0][2] = b[i - 1 + sz]; m[0][3] = b[i - 1 + sy + sz];
        q[0][0] = b[i + 2]; q[0][1] = b[i + 2 + sy]; q[0][2] = b[i + 2 + sz]; q[0][3] = b[i + 2 + sy + sz];
        // y: (x, z) = (0,0), (1,0), (0,1), (1,1)
        m[1][0] = b[i - sy]; m[1][1] = b[i - sy + 1]; m[1][2] = b[i - sy + sz]; m[1][3] = b[i - sy + sz + 1];
        q[1][0] = b[i + 2 * sy]; q[1][1] = b[i + 2 * sy + 1]; q[1][2] = b[i + 2 * sy + sz];
        q[1][3] = b[i + 2 * sy + sz + 1];
        // z: (x, y) = (0,0), (1,0), (0,1), (1,1)
        m[2][0] = b[i - sz]; m[2][1] = b[i - sz + 1]; m[2][2] = b[i - sz + sy]; m[2][3] = b[i - sz + sy + 1];
        q[2][0] = b[i + 2 * sz]; q[2][1] = b[i + 2 * sz + 1]; q[2][2] = b[i + 2 * sz + sy];
        q[2][3] = b[i + 2 * sz + sy + 1];
    } else {
        m[0][0] = rd.gvox(x - 1, y, z); m[0][1] = rd.gvox(x - 1, y + 1, z);
        m[0][2] = rd.gvox(x - 1, y, z + 1); m[0][3] = rd.gvox(x - 1, y + 1, z + 1);
        q[0][0] = rd.gvox(x + 2, y, z); q[0][1] = rd.gvox(x + 2, y + 1, z);
        q[0][2] = rd.gvox(x + 2, y, z + 1); q[0][3] = rd.gvox(x + 2, y + 1, z + 1);
        m[1][0] = rd.gvox(x, y - 1, z); m[1][1] = rd.gvox(x + 1, y - 1, z);
        m[1][2] = rd.gvox(x, y - 1, z + 1); m[1][3] = rd.gvox(x + 1, y - 1, z + 1);
        q[1][0] = rd.gvox(x, y + 2, z); q[1][1] = rd.gvox(x + 1, y + 2, z);
        q[1][2] = rd.gvox(x, y + 2, z + 1); q[1][3] = rd.gvox(x + 1, y + 2, z + 1);
        m[2][0] = rd.gvox(x, y, z - 1); m[2][1] = rd.gvox(x + 1, y, z - 1);
        m[2][2] = rd.gvox(x, y + 1, z - 1); m[2][3] = rd.gvox(x + 1, y + 1, z - 1);
        q[2][0] = rd.gvox(x, y, z + 2); q[2][1] = rd.gvox(x + 1, y, z + 2);
        q[2][2] = rd.gvox(x, y + 1, z + 2); q[2][3] = rd.gvox(x + 1, y + 1, z + 2);
    }
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (float)v8[k];
    const float px = p[0], py = p[1], pz = p[2];
    const float qx = 1.f - px, qy = 1.f - py, qz = 1.f - pz;
    auto dot4 = [](float a0, float a1, float a2, float a3, const float w[4]) {
        return __builtin_fmaf(a3, w[3], __builtin_fmaf(a2, w[2], __builtin_fmaf(a1, w[1], a0 * w[0])));
    };
    auto diff = [](float qa, float pa, float bm, float b0, float b1, float bp) {
        return __builtin_fmaf(-pa, bp, __builtin_fmaf(-qa, b1, __builtin_fmaf(pa, b0, qa * bm)));
    };
    // corner order: 0 x0y0z0, 1 x1y0z0, 2 x1y1z0, 3 x0y1z0, 4 x0y0z1, 5 x1y0z1, 6 x1y1z1, 7 x0y1z1
    {
        const float w[4] = {qy * qz, py * qz, qy * pz, py * pz};
        const float bm = dot4((float)m[0][0], (float)m[0][1], (float)m[0][2], (float)m[0][3], w);
        const float b0 = dot4(v[0], v[3], v[4], v[7], w);
        const float b1 = dot4(v[1], v[2], v[5], v[6], w);
        const float bp = dot4((float)q[0][0], (float)q[0][1], (float)q[0][2], (float)q[0][3], w);
        gr[0] = diff(qx, px, bm, b0, b1, bp) * a.grad_raw[0];
    }
    {
        const float w[4] = {qx * qz, px * qz, qx * pz, px * pz};
        const float bm = dot4((float)m[1][0], (float)m[1][1], (float)m[1][2], (float)m[1][3], w);
        const float b0 = dot4(v[0], v[1], v[4], v[5], w);
        const float b1 = dot4(v[3], v[2], v[7], v[6], w);
        const float bp = dot4((float)q[1][0], (float)q[1][1], (float)q[1][2], (float)q[1][3], w);
        gr[1] = diff(qy, py, bm, b0, b1, bp) * a.grad_raw[1];
    }
    {
        const float w[4] = {qx * qy, px * qy, qx * py, px * py};
        const float bm = dot4((float)m[2][0], (float)m[2][1], (float)m[2][2], (float)m[2][3], w);
        const float b0 = dot4(v[0], v[1], v[3], v[2], w);
        const float b1 = dot4(v[4], v[5], v[7], v[6], w);
        const float bp = dot4((float)q[2][0], (float)q[2][1], (float)q[2][2], (float)q[2][3], w);
        gr[2] = diff(qz, pz, bm, b0, b1, bp) * a.grad_raw[2];
    }
    return true;
}

// Per-ray state of the slab march (kept in registers across rounds).
struct SlabRay {
    float first[3], dir[3];
    float tend, t;
    float r, g, b, A;
    int loop;
    int px, py;
    uint32_t cur_brick;
    float t_lim;
    bool cur_occ;
    bool active;       // a ray that is not finished
};

// March one ray from its state while its next data sample lies in slab
// `slab` along axis d (-1: no slab is staged, only empty bricks are crossed;
// kAnySlab: every sample, to the ray's end).  Returns the key of the slab its next data sample needs, or
// kNoSlab when the ray has finished.  The sample loop is the cell kernel's
// with one sample at a time.
template <typename T, int SHADE, bool COUNT>
__device__ __forceinline__ int slab_march(const MarchArgs& a, const SlabReader<T>& rd, const uint32_t* s_tf,
                                          const float* s_alpha, const uint32_t* s_occ, int d, int slab,
                                          SlabRay& ray, RayCounters& n) {
    const uint32_t bl = a.brick_log2;
    const uint32_t nbx = a.nbricks[0], nbxy = a.nbricks[0] * a.nbricks[1];
    const int cap = a.loop_cap;
    while (ray.active && ray.loop < cap) {
        const float s[3] = {ray.first[0] + ray.t * ray.dir[0], ray.first[1] + ray.t * ray.dir[1],
                            ray.first[2] + ray.t * ray.dir[2]};
        float vp[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
        const uint32_t brick = (((uint32_t)vp[2] >> bl) * nbxy) + (((uint32_t)vp[1] >> bl) * nbx) +
                               ((uint32_t)vp[0] >> bl);
        if (brick != ray.cur_brick) {
            ray.cur_brick = brick;
            ray.cur_occ = s_occ ? ((s_occ[brick >> 5] >> (brick & 31)) & 1u) != 0 : a.brick_occ[brick] != 0;
            if (!ray.cur_occ) ray.t_lim = empty_brick_limit(a, ray.first, ray.dir, vp, bl);
        }
        if (!ray.cur_occ) {
            // every sample at t' < t_lim lies in this empty brick
            ++n.samples;
            if (COUNT) ++n.slots;
            {
                float dd = 0.0f;
                const int j = closed_form_steps(ray.t, a.step, fminf(ray.t_lim, ray.tend), cap - ray.loop, dd);
                if (j > 0) {
                    ray.t = ray.t + (float)j * dd;
                    ray.loop += j;
                    n.samples += (unsigned long long)j;
                    if (COUNT) n.slots += (unsigned long long)j;
                }
            }
            float tn = ray.t + a.step;
            while (tn < ray.t_lim && !(tn > ray.tend) && ray.loop + 1 < cap) {
                ray.t = tn;
                ++ray.loop;
                ++n.samples;
                if (COUNT) ++n.slots;
                tn = ray.t + a.step;
            }
            ++ray.loop;
            ray.t = tn;
            ray.active = !(ray.t > ray.tend);
            continue;
        }
        uint32_t lo[3];
        float p[3];
        slab_corner(a, s, lo, p);
        const int sd = (int)(lo[d] >> kSlabLog2);
        if (slab != kAnySlab && sd != slab) return sd;   // the next data sample is in another slab
        ++ray.loop;
        ++n.samples;
        ++n.fetched;
        uint32_t v8[8];
        rd.corners8(lo, v8);
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = VoxNorm<T>::apply(v8[k]);
        const float inten = trilerp(v, p);
        const uint32_t c8 = s_tf[tf_index(inten, a)];
        const float ca = s_alpha[c8 >> 24];
        bool shade_now = false;
        if (ca > 0.0f) {
            float col[3] = {byte_norm(c8 & 0xffu), byte_norm((c8 >> 8) & 0xffu), byte_norm((c8 >> 16) & 0xffu)};
            if (SHADE) {
                shade_now = true;
                float gr[3];
                if (ray.t == 0.0f) clip_border_gradient(a, ray.px, ray.py, s, ray.dir, gr);
                else if (SHADE == 2) slab_gradient6<T, false>(rd, a, s, gr);
                else if (!slab_gradient_interior<T>(rd, a, s, lo, p, v8, gr)) slab_gradient6<T, true>(rd, a, s, gr);
                float pos[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) pos[i] = (((s[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                if (SHADE == 2) shade_model(a, gr, pos, col);
                else phong_fast(a, gr, pos, col);
                ++n.shaded;
            }
            const float wgt = (1.0f - ray.A) * ca;
            ray.r = ray.r + wgt * col[0];
            ray.g = ray.g + wgt * col[1];
            ray.b = ray.b + wgt * col[2];
            ray.A = ray.A + (1.0f - ray.A) * ca;
        }
        if (COUNT) {
            ++n.slots;
            n.shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
        }
        bool done = false;
        if (ray.A >= kErtOpacity) {
            ray.A = 1.0f;
            done = true;
        }
        ray.t = ray.t + a.step;
        ray.active = !(done || (ray.t > ray.tend));
    }
    ray.active = false;
    return kNoSlab;
}

}  // namespace hipdvr
