// learned_ray.h -- the learned marcher's ray loop, shared by the per-frame
// learned kernel (learned_march.hip) and the batched launch (batch.hip).
#pragma once

#include "learned.h"
#include "march_cell.h"
#include "march_common.h"

namespace hipdvr {

// Ray parameter where the ray leaves the axis-aligned voxel box [lo, hi)
// (unclamped voxel coordinates; first/last bricks of an axis extend to
// infinity), plus a small push so the next sample is past the boundary.
__device__ __forceinline__ float box_exit(const MarchArgs& a, const float first[3], const float dir[3],
                                          const uint32_t b[3], uint32_t B, const uint32_t nb[3]) {
    float tl = INFINITY;
    for (int i = 0; i < 3; ++i) {
        if (dir[i] > 0.0f && b[i] + 1 < nb[i]) {
            const float sb = ((float)((b[i] + 1) * B) + 0.5f) / a.fdims[i];
            tl = fminf(tl, (sb - first[i]) / dir[i]);
        } else if (dir[i] < 0.0f && b[i] > 0) {
            const float sb = ((float)(b[i] * B) + 0.5f) / a.fdims[i];
            tl = fminf(tl, (sb - first[i]) / dir[i]);
        }
    }
    return tl;
}

// box_exit with the divisions taken once per ray: the reciprocals of the
// direction and of the volume's dims.  The brick faces are then
// ((k*B + 0.5) * (1/dim) - first) * (1/dir), a few ulps from box_exit.  A
// wavefront enters its rays' next bricks at different samples, so it runs
// this on most passes of the march loop; the learned march and the trace
// that labels its training rows both use it.
struct BoxExit {
    float inv_dir[3], inv_dim[3];
};

__device__ __forceinline__ BoxExit box_exit_setup(const MarchArgs& a, const float dir[3]) {
    BoxExit x;
    for (int i = 0; i < 3; ++i) {
        x.inv_dir[i] = 1.0f / dir[i];
        x.inv_dim[i] = 1.0f / a.fdims[i];
    }
    return x;
}

__device__ __forceinline__ float box_exit_fast(const BoxExit& x, const float first[3], const float dir[3],
                                               const uint32_t b[3], uint32_t B, const uint32_t nb[3]) {
    float tl = INFINITY;
    for (int i = 0; i < 3; ++i) {
        if (dir[i] > 0.0f && b[i] + 1 < nb[i]) {
            const float sb = ((float)((b[i] + 1) * B) + 0.5f) * x.inv_dim[i];
            tl = fminf(tl, (sb - first[i]) * x.inv_dir[i]);
        } else if (dir[i] < 0.0f && b[i] > 0) {
            const float sb = ((float)(b[i] * B) + 0.5f) * x.inv_dim[i];
            tl = fminf(tl, (sb - first[i]) * x.inv_dir[i]);
        }
    }
    return tl;
}
constexpr bool kLearnedFastExit = true;

// ---------------------------------------------------------------------------
// the learned marcher: the cell-record march (march_cell.h) with the stepping
// table in place of exact empty-space skipping.  Per 16^3 brick and ray
// octant the table says skip (jump to the first step-grid point past the
// brick) or march with a step multiple m in {1, 2, 4, 8}; the last step in a
// brick is cut so the next brick starts on the base grid, and each sample's
// opacity correction is the one of the step it stands for.
// ---------------------------------------------------------------------------
// Opacity of a step m * step (m = 1 .. 8) from the corrected opacity a1 of
// one step: 1 - (1 - a1)^m, since (1 - a)^(m*e) = ((1 - a)^e)^m.  (A table
// per multiple would cost 7 KiB of LDS per workgroup, a quarter of the
// kernel's occupancy.)
__device__ __forceinline__ float alpha_for_multiple(float a1, int m) {
    if (m == 1) return a1;
    const float q = 1.0f - a1, q2 = q * q, q4 = q2 * q2;
    if (m >= 8) return 1.0f - q4 * q4;
    const float qm = ((m & 1) ? q : 1.0f) * ((m & 2) ? q2 : 1.0f) * ((m & 4) ? q4 : 1.0f);
    return 1.0f - qm;
}

// One ray of the learned march from entry point e to exit point l (neither
// the background): returns the composited RGBA.  K samples fetched together
// per ray: two when shaded (as the plain march; 1069 against 1018 frames/s
// with four on 512^3 / 1024^2), else four.
template <typename T, bool WIDE, bool PHONG>
__device__ __forceinline__ float4 learned_ray(const LearnedArgs& la, const CellFetch<T, WIDE>& f,
                                              const uint32_t* s_tf, const float* s_alpha, int px, int py, float4 e,
                                              float4 l, unsigned long long& n_samples,
                                              unsigned long long& n_shaded) {
    constexpr int K = PHONG ? 2 : 4;
    typedef typename CellFetch<T, WIDE>::rec_t rec_t;
    const MarchArgs& a = la.m;
    const float first[3] = {e.x, e.y, e.z}, last[3] = {l.x, l.y, l.z};
    float dir[3];
    const float tend = ray_setup(first, last, a.step, dir);
    const uint32_t oct = (dir[0] < 0.f ? 1u : 0u) | (dir[1] < 0.f ? 2u : 0u) | (dir[2] < 0.f ? 4u : 0u);
    const BoxExit bx = box_exit_setup(a, dir);
    float t = 0.f, r = 0.f, g = 0.f, bb = 0.f, A = 0.f;
    int guard = 0;
    bool finished = false;
    uint32_t cur = 0xffffffffu;   // brick of the cached table entry
    float t_exit = 0.f;
    int mmax = 1;
    bool skip = false;
    while (!finished && !(t > tend) && guard < a.loop_cap) {
        float s[3], vp[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = first[i] + t * dir[i];
            vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
        }
        uint32_t bc[3];
        for (int i = 0; i < 3; ++i) bc[i] = (uint32_t)vp[i] >> kLBrickLog2;
        const uint32_t brick = (bc[2] * la.nb16[1] + bc[1]) * la.nb16[0] + bc[0];
        if (brick != cur) {
            cur = brick;
            const uint8_t ent = la.table[brick * 8 + oct];
            t_exit = kLearnedFastExit ? box_exit_fast(bx, first, dir, bc, kLBrick, la.nb16)
                                      : box_exit(a, first, dir, bc, kLBrick, la.nb16);
            skip = (ent & 0x80) != 0;
            mmax = 1 << (ent & 3);
        }
        if (skip) {
            // skip to the first step-grid point past the brick (a point that
            // rounding leaves inside it is stepped on by the next pass)
            const float k = floorf((t_exit - t) / a.step);
            t = t + (fmaxf(k, 0.f) + 1.f) * a.step;
            ++guard;
            continue;
        }
        // up to K samples of this brick, their loads in flight together
        float tk[K], pk[K][3];
        int mk[K];
        rec_t rk[K];
        typename CellFetch<T, WIDE>::off_t ik[K];
        bool valid[K];
        float tt = t;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            valid[k] = (k == 0) || (valid[k - 1] && !(tt > tend) && tt < t_exit);
            tk[k] = tt;
            int m = mmax;
            if (m > 1 && tt + (float)m * a.step > t_exit) {
                const float c = ceilf((t_exit - tt) / a.step);
                m = c < 1.f ? 1 : (c < (float)m ? (int)c : m);
            }
            mk[k] = m;
            if (valid[k]) {
                const float sk[3] = {first[0] + tt * dir[0], first[1] + tt * dir[1], first[2] + tt * dir[2]};
                ik[k] = cell_of<T, WIDE>(f, a, sk, pk[k]);
                rk[k] = f.load(ik[k]);
            }
            tt = tt + (float)m * a.step;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!valid[k] || finished) break;
            ++n_samples;
            ++guard;
            float v[8];
            CellRecord<T>::unpack(rk[k], v);
            const float inten = trilerp(v, pk[k]);
            const uint32_t c8 = s_tf[tf_index(inten, a)];
            const float ca = alpha_for_multiple(s_alpha[c8 >> 24], mk[k]);
            if (ca > 0.f) {
                float col[3] = {byte_norm(c8 & 0xffu), byte_norm((c8 >> 8) & 0xffu),
                                byte_norm((c8 >> 16) & 0xffu)};
                if (PHONG) {
                    const float t0 = tk[k];
                    const float ss[3] = {first[0] + t0 * dir[0], first[1] + t0 * dir[1],
                                         first[2] + t0 * dir[2]};
                    float gr[3];
                    if (t0 == 0.0f) clip_border_gradient(a, px, py, ss, dir, gr);
                    else if (!cell_gradient_interior<T, WIDE>(f, a, ss, rk[k], ik[k], pk[k], gr))
                        cell_gradient<T, WIDE, true>(f, a, ss, gr);
                    float pos[3];
                    for (int i = 0; i < 3; ++i)
                        pos[i] = (((ss[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                    phong_fast(a, gr, pos, col);
                    ++n_shaded;
                }
                const float w = (1.f - A) * ca;
                r += w * col[0];
                g += w * col[1];
                bb += w * col[2];
                A += w;
            }
            if (A >= kErtOpacity) {
                A = 1.f;
                finished = true;
            }
            t = tk[k] + (float)mk[k] * a.step;
        }
    }
    return make_float4(r, g, bb, A);
}

}  // namespace hipdvr
