// march_ref.hip -- reference-layout ray march: linear (x-fastest) volume, one
// lane per ray, 8x8 pixel tile per wavefront, 16x16 pixels per workgroup.
//
// Restates CPURaycaster::directRendering
// (modules/base/processors/render/cpuraycaster.cpp:210-326) sample by sample:
// same sample positions (t accumulated in float), same trilinear order, same
// TF indexing and opacity correction, same ERT.  The TF (RGBA8, <= 4096
// entries) and the 256-entry opacity-correction table sit in LDS.
#include "march_common.h"
#include "kernels.h"

namespace hipdvr {

template <typename T>
struct LinearFetch {
    const T* p;
    uint32_t X;
    uint64_t XY;
    __device__ __forceinline__ void fetch8(const Corners& c, float v[8]) const {
        uint64_t base = (uint64_t)c.lo[2] * XY + (uint64_t)c.lo[1] * X + c.lo[0];
        uint64_t dx = c.d[0], dy = (uint64_t)c.d[1] * X, dz = (uint64_t)c.d[2] * XY;
        uint32_t r0 = p[base];
        uint32_t r1 = p[base + dx];
        uint32_t r2 = p[base + dx + dy];
        uint32_t r3 = p[base + dy];
        uint32_t r4 = p[base + dz];
        uint32_t r5 = p[base + dx + dz];
        uint32_t r6 = p[base + dx + dy + dz];
        uint32_t r7 = p[base + dy + dz];
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
        uint32_t x = (uint32_t)iround_tgt(vp[0]);
        uint32_t y = (uint32_t)iround_tgt(vp[1]);
        uint32_t z = (uint32_t)iround_tgt(vp[2]);
        return VoxNorm<T>::apply(p[(uint64_t)z * XY + (uint64_t)y * X + x]);
    }
};

template <typename T, int FILTER>
__device__ __forceinline__ float reconstruct(const LinearFetch<T>& f, const MarchArgs& a,
                                             const float s[3]) {
    float vp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
    if (FILTER == 0) return f.nearest(vp);
    Corners c;
    corners(vp, a.dims, c);
    float v[8];
    f.fetch8(c, v);
    return trilerp(v, c.p);
}

// calcGradientA, mod_gradients.frag:202-214: central differences at +-1/dim
// in texture space (down-gradient), times spacingRCP*0.5 and rwmScale.
template <typename T, int FILTER>
__device__ __forceinline__ void grad_central(const LinearFetch<T>& f, const MarchArgs& a, const float s[3],
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
        vals[k] = reconstruct<T, FILTER>(f, a, q);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) gr[i] = ((vals[3 + i] - vals[i]) * a.grad_scale[i]) * a.rwm_scale;
}

// calcGradientSobel, mod_gradients.frag:224-286.  V[n] is the shader's vNN:
// n-1 = xg*9 + zg*3 + yg, xg = 0, 1, 2 for x = +1, 0, -1 (in units of 1/dim),
// zg likewise for z, yg = 0, 1, 2 for y = -1, 0, +1.  The sums run left to
// right as written; the factors 2 and 4 scale exactly.
template <typename T, int FILTER>
__device__ __noinline__ void grad_sobel(const LinearFetch<T>& f, const MarchArgs& a, const float s[3], float gr[3]) {
    float V[28];
    V[0] = V[14] = 0.f;
#pragma unroll
    for (int n = 1; n <= 27; ++n) {
        if (n == 14) continue;
        const int k = n - 1, sx = 1 - k / 9, sz = 1 - (k / 3) % 3, sy = k % 3 - 1;
        float q[3];
        q[0] = s[0] + (sx > 0 ? a.grad_off[0] : (sx < 0 ? -a.grad_off[0] : 0.0f));
        q[1] = s[1] + (sy > 0 ? a.grad_off[1] : (sy < 0 ? -a.grad_off[1] : 0.0f));
        q[2] = s[2] + (sz > 0 ? a.grad_off[2] : (sz < 0 ? -a.grad_off[2] : 0.0f));
        V[n] = reconstruct<T, FILTER>(f, a, q);
    }
    const float gx = V[1] + 2.0f * V[2] + V[3] - V[19] - 2.0f * V[20] - V[21]
                   + 2.0f * V[4] + 4.0f * V[5] + 2.0f * V[6] - 2.0f * V[22] - 4.0f * V[23] - 2.0f * V[24]
                   + V[7] + 2.0f * V[8] + V[9] - V[25] - 2.0f * V[26] - V[27];
    const float gy = -V[1] + V[3] - 2.0f * V[10] + 2.0f * V[12] - V[19] + V[21]
                   - 2.0f * V[4] + 2.0f * V[6] - 4.0f * V[13] + 4.0f * V[15] - 2.0f * V[22] + 2.0f * V[24]
                   - V[7] + V[9] - 2.0f * V[16] + 2.0f * V[18] - V[25] + V[27];
    const float gz = V[1] + 2.0f * V[2] + V[3] + 2.0f * V[10] + 4.0f * V[11] + 2.0f * V[12] + V[19]
                   + 2.0f * V[20] + V[21] - V[7] - 2.0f * V[8] - V[9] - 2.0f * V[16] - 4.0f * V[17]
                   - 2.0f * V[18] - V[25] - 2.0f * V[26] - V[27];
    const float g[3] = {-gx, -gy, -gz};
#pragma unroll
    for (int i = 0; i < 3; ++i) gr[i] = ((g[i] * a.grad_scale[i]) / 16.0f) * a.rwm_scale;
}

// calcGradientFiltered, mod_gradients.frag:298-314: central gradients at the
// sample and its eight diagonal neighbours blended by GLSL mix
// (x*(1-a) + y*a).
template <typename T, int FILTER>
__device__ __noinline__ void grad_filtered(const LinearFetch<T>& f, const MarchArgs& a, const float s[3],
                                           float gr[3]) {
    float G[9][3];
    grad_central<T, FILTER>(f, a, s, G[0]);
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        // diagonals in the shader's order: ---, +++, -+-, +-+, --+, ++-, -++, +--
        const uint32_t neg = 0xC63547u >> (3 * k);   // bit i set: axis i steps by -1/dim
        float q[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = s[i] + (((neg >> i) & 1u) ? -a.grad_off[i] : a.grad_off[i]);
        float g[3];
        grad_central<T, FILTER>(f, a, q, g);
#pragma unroll
        for (int i = 0; i < 3; ++i) G[k + 1][i] = g[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float m12 = G[1][i] * 0.5f + G[2][i] * 0.5f, m34 = G[3][i] * 0.5f + G[4][i] * 0.5f;
        const float m56 = G[5][i] * 0.5f + G[6][i] * 0.5f, m78 = G[7][i] * 0.5f + G[8][i] * 0.5f;
        const float mix0 = m12 * 0.5f + m34 * 0.5f, mix1 = m56 * 0.5f + m78 * 0.5f;
        gr[i] = G[0][i] * 0.25f + (mix0 * 0.5f + mix1 * 0.5f) * 0.75f;
    }
}

// CALC_GRADIENT as the "gradient" property configures it (volumeraycaster.cpp:197-207).
template <typename T, int FILTER>
__device__ __forceinline__ void sample_gradient(const LinearFetch<T>& f, const MarchArgs& a, const float s[3],
                                                float inten, float gr[3]) {
    if (a.gradient == kGradCentral) {
        grad_central<T, FILTER>(f, a, s, gr);
    } else if (a.gradient == kGradForward) {
        // calcGradientAFD, mod_gradients.frag:116-126: (v(s) - v(s + 1/dim)) * spacingRCP * rwmScale
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            float q[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) q[i] = s[i] + (i == axis ? a.grad_off[i] : 0.0f);
            gr[axis] = ((inten - reconstruct<T, FILTER>(f, a, q)) * a.spacing_rcp[axis]) * a.rwm_scale;
        }
    } else if (a.gradient == kGradNone) {
        // (voxel.xyz - 0.5)*2 on the sample of a one-channel texture swizzled to
        // (R, 0, 0, R) (volumegl.cpp:367), real-world mapping on every component
        // (mod_sampler3d.frag:90-101)
        const float yz = (0.0f * a.rwm_scale) + a.rwm_offset;
        gr[0] = (((inten * a.rwm_scale) + a.rwm_offset) - 0.5f) * 2.0f;
        gr[1] = (yz - 0.5f) * 2.0f;
        gr[2] = (yz - 0.5f) * 2.0f;
    } else if (a.gradient == kGradFiltered) {
        grad_filtered<T, FILTER>(f, a, s, gr);
    } else {
        grad_sobel<T, FILTER>(f, a, s, gr);
    }
}

// Channel c of gradient-volume voxel `vox`, normalized (getVoxelNormalized(pos,
// channel), volumeatomic.h:531-536; float channels as stored).
__device__ __forceinline__ float gvol_norm(const MarchArgs& a, uint64_t vox, int c) {
    const uint64_t i = vox * (uint32_t)a.gvol_channels + (uint32_t)c;
    if (a.gvol_dtype == 0) return VoxNorm<uint8_t>::apply(((const uint8_t*)a.gvol)[i]);
    if (a.gvol_dtype == 1) return VoxNorm<uint16_t>::apply(((const uint16_t*)a.gvol)[i]);
    return ((const float*)a.gvol)[i];
}

// The gradient-volume sample of the 2D TF path (cpuraycaster.cpp:270-284):
// channels 0..2 at the rounded voxel position, or trilinear per channel.
template <int FILTER>
__device__ __forceinline__ void gvol_sample(const MarchArgs& a, const float vp[3], float out[3]) {
    const uint32_t X = a.dims[0];
    const uint64_t XY = (uint64_t)a.dims[0] * a.dims[1];
    if (FILTER == 0) {
        const uint64_t v = (uint64_t)(uint32_t)iround_tgt(vp[2]) * XY + (uint64_t)(uint32_t)iround_tgt(vp[1]) * X +
                           (uint32_t)iround_tgt(vp[0]);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = gvol_norm(a, v, c);
        return;
    }
    Corners cr;
    corners(vp, a.dims, cr);
    const uint64_t base = (uint64_t)cr.lo[2] * XY + (uint64_t)cr.lo[1] * X + cr.lo[0];
    const uint64_t dx = cr.d[0], dy = (uint64_t)cr.d[1] * X, dz = (uint64_t)cr.d[2] * XY;
    const uint64_t at[8] = {base, base + dx, base + dx + dy, base + dy,
                            base + dz, base + dx + dz, base + dx + dy + dz, base + dy + dz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = gvol_norm(a, at[k], c);
        out[c] = trilerp(v, cr.p);
    }
}

template <typename T, int FILTER, int CLASSIF, bool PHONG, bool COUNT>
__global__ __launch_bounds__(256) void march_ref_kernel(const MarchArgs a) {
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    const int tid = threadIdx.x;
    if (CLASSIF == 0) {
        for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
        s_alpha[tid] = a.alpha_corr[tid];
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = (px < a.w) && (py < a.h);

    LinearFetch<T> f;
    f.p = (const T*)a.vox;
    f.X = a.dims[0];
    f.XY = (uint64_t)a.dims[0] * a.dims[1];

    unsigned long long n_slots = 0, n_shade_slots = 0, n_samples = 0, n_shaded = 0, n_rays = 0;
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
            float t = 0.0f, r = 0.f, g = 0.f, b = 0.f, A = 0.f, last_x = 0.f;
            bool finished = false;
            for (int loop = 0; !finished && loop < cap; ++loop) {
                ++n_samples;
                bool shade_now = false;
                const float s[3] = {first[0] + t * dir[0], first[1] + t * dir[1],
                                    first[2] + t * dir[2]};
                const float inten = reconstruct<T, FILTER>(f, a, s);
                float cr, cg, cb, ca;
                if (CLASSIF == 1) {
                    const float x = norm_rw(inten, a);
                    const int n = a.preint_n;
                    const float4 c = a.preint[(size_t)preint_index(x, n) * n + preint_index(last_x, n)];
                    cr = c.x; cg = c.y; cb = c.z; ca = c.w;
                    last_x = x;
                } else if (CLASSIF == 2) {
                    // cpuraycaster.cpp:270-290 with apply2DTF :334-338; the table
                    // holds the opacity-corrected alpha
                    float vp[3], gv[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) vp[i] = tex_to_voxel(s[i], a.fdims[i], a.fdm1[i]);
                    gvol_sample<FILTER>(a, vp, gv);
                    float lsq = 0.0f;
                    lsq = lsq + gv[0] * gv[0];
                    lsq = lsq + gv[1] * gv[1];
                    lsq = lsq + gv[2] * gv[2];
                    const float rw = (inten * a.rwm_scale) + a.rwm_offset;
                    const float xi = rw_to_norm(rw, a.tf2d_dom[0], a.tf2d_dom[1]);
                    const float xg = rw_to_norm(sqrtf(lsq), a.tf2d_dom[2], a.tf2d_dom[3]);
                    // truncation as size_t(x*(W-1)); x is in [0, 1] (a NaN
                    // magnitude from float gradient data selects texel 0)
                    const float fx = xi * (float)(a.tf2d_w - 1), fy = xg * (float)(a.tf2d_h - 1);
                    const uint32_t ix = fx > 0.0f ? min((uint32_t)fx, (uint32_t)a.tf2d_w - 1u) : 0u;
                    const uint32_t iy = fy > 0.0f ? min((uint32_t)fy, (uint32_t)a.tf2d_h - 1u) : 0u;
                    const float4 c = a.tf2d[(size_t)iy * (uint32_t)a.tf2d_w + ix];
                    cr = c.x; cg = c.y; cb = c.z; ca = c.w;
                } else {
                    const uint32_t c8 = s_tf[tf_index(inten, a)];
                    ca = s_alpha[c8 >> 24];
                    cr = byte_norm(c8 & 0xffu);
                    cg = byte_norm((c8 >> 8) & 0xffu);
                    cb = byte_norm((c8 >> 16) & 0xffu);
                }
                if (ca > 0.0f) {
                    shade_now = PHONG;
                    if (PHONG) {
                        float gr[3];
                        if (t == 0.0f) clip_border_gradient(a, px, py, s, dir, gr);
                        else sample_gradient<T, FILTER>(f, a, s, inten, gr);
                        float pos[3], c[3] = {cr, cg, cb};
#pragma unroll
                        for (int i = 0; i < 3; ++i)
                            pos[i] = (((s[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                        shade_model(a, gr, pos, c);
                        cr = c[0]; cg = c[1]; cb = c[2];
                        ++n_shaded;
                    }
                    const float wgt = (1.0f - A) * ca;
                    r = r + wgt * cr;
                    g = g + wgt * cg;
                    b = b + wgt * cb;
                    A = A + (1.0f - A) * ca;
                }
                if (A >= kErtOpacity) {
                    A = 1.0f;
                    finished = true;
                }
                if (COUNT) {
                    ++n_slots;
                    n_shade_slots += (__ballot(shade_now) != 0ull) ? 1 : 0;
                }
                t += a.step;
                finished = finished || (t > tend);
            }
            res = make_float4(r, g, b, A);
        }
        a.out[pix] = res;
    }
    if (COUNT) {
        wave_add(a.counters + 0, n_rays);
        wave_add(a.counters + 1, n_samples);
        wave_add(a.counters + 2, n_samples);
        wave_add(a.counters + 3, n_shaded);
        wave_add_max64(a.counters + 4, n_slots);
        wave_add_max64(a.counters + 5, n_shade_slots);
    }
}

// compositeMIDAhelper, mod_compositing.frag:112-129 (maxIntensity read only).
__device__ __forceinline__ void mida_helper(float res[4], float voxel, const float col[4], float max_int,
                                            float gamma) {
    float delta = 0.0f;
    if (voxel > max_int) delta = voxel - max_int;
    const float beta = 1.0f - delta * (1.0f + gamma);
    const float k = (1.0f - beta * res[3]) * col[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) res[i] = beta * res[i] + k * col[i];
    res[3] = beta * res[3] + (1.0f - beta * res[3]) * col[3];
}

// Compositing modes other than DVR with TF classification: the GLSL
// compositing functions (mod_compositing.frag:85-213) applied where the TF
// alpha is > 0, after shading (rc_singlevolume.frag:222-236), on the
// CPURaycaster sample sequence; early termination only for ISO, FHP and FHN
// (volumeraycaster.cpp:225-233).
template <typename T, int FILTER, bool PHONG>
__global__ __launch_bounds__(256) void march_modes_kernel(const MarchArgs a) {
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = px < a.w && py < a.h;
    unsigned long long n_rays = 0, n_samples = 0, n_shaded = 0;
    LinearFetch<T> f;
    f.p = (const T*)a.vox;
    f.X = a.dims[0];
    f.XY = (uint64_t)a.dims[0] * a.dims[1];
    const int mode = a.compositing;
    const bool ert = mode == kCompISO || mode == kCompFHP || mode == kCompFHN;
    const size_t pix = (size_t)py * a.w + px;
    const float4 e = inside ? a.entry[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 l = inside ? a.exit[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    float out[4] = {0.f, 0.f, 0.f, 0.f};
    if (inside && !is_background(e, l)) {
        n_rays = 1;
        const float first[3] = {e.x, e.y, e.z};
        const float last[3] = {l.x, l.y, l.z};
        float dir[3];
        const float tend = ray_setup(first, last, a.step, dir);
        float t = 0.0f, max_int = 0.0f;
        bool finished = false;
        for (int loop = 0; !finished && loop < a.loop_cap; ++loop) {
            ++n_samples;
            const float s[3] = {first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2]};
            const float inten = reconstruct<T, FILTER>(f, a, s);
            const float rw = (inten * a.rwm_scale) + a.rwm_offset;
            const uint32_t c8 = s_tf[tf_index(inten, a)];
            float col[4] = {byte_norm(c8 & 0xffu), byte_norm((c8 >> 8) & 0xffu), byte_norm((c8 >> 16) & 0xffu),
                            byte_norm(c8 >> 24)};
            if (col[3] > 0.0f) {
                float gr[3] = {0.f, 0.f, 0.f};
                if (PHONG || mode == kCompFHN) {
                    if (t == 0.0f) clip_border_gradient(a, px, py, s, dir, gr);
                    else sample_gradient<T, FILTER>(f, a, s, inten, gr);
                }
                if (PHONG) {
                    float pos[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) pos[i] = (((s[i] * a.fdims[i]) - 0.5f) * a.spacing[i]) + a.offset[i];
                    shade_model(a, gr, pos, col);
                    ++n_shaded;
                }
                if (mode == kCompMIP) {
                    if (rw > max_int) {
                        for (int i = 0; i < 4; ++i) out[i] = col[i];
                        max_int = rw;
                    }
                } else if (mode == kCompMOP) {
                    if (col[3] > out[3])
                        for (int i = 0; i < 4; ++i) out[i] = col[i];
                } else if (mode == kCompMIDA) {
                    col[3] = s_alpha[c8 >> 24];
                    if (a.gamma <= 0.0f) {
                        mida_helper(out, rw, col, max_int, a.gamma);
                    } else {
                        float mip[4], mida[4];
                        for (int i = 0; i < 4; ++i) mip[i] = mida[i] = out[i];
                        if (rw > max_int) {
                            for (int i = 0; i < 4; ++i) mip[i] = col[i];
                            max_int = rw;
                        }
                        mida_helper(mida, rw, col, max_int, 0.0f);
                        for (int i = 0; i < 4; ++i) out[i] = a.gamma * mip[i] + (1.0f - a.gamma) * mida[i];
                    }
                } else if (mode == kCompISO) {
                    if (col[3] >= a.iso_value - 0.02f && col[3] <= a.iso_value + 0.02f) {
                        out[0] = col[0];
                        out[1] = col[1];
                        out[2] = col[2];
                        out[3] = 1.0f;
                    }
                } else if (mode == kCompFHP) {
                    if (out[0] == 0.f && out[1] == 0.f && out[2] == 0.f) {
                        out[0] = s[0];
                        out[1] = s[1];
                        out[2] = s[2];
                        out[3] = 1.0f;
                    }
                } else if (mode == kCompFHN) {
                    if (out[0] == 0.f && out[1] == 0.f && out[2] == 0.f) {
                        const float l2 = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
                        const float rl = l2 > 0.f ? 1.0f / sqrtf(l2) : 0.f;
                        for (int i = 0; i < 3; ++i) out[i] = ((gr[i] * rl) / 2.0f) + 0.5f;
                        out[3] = 1.0f;
                    }
                }
            }
            if (ert && out[3] >= kErtOpacity) {
                out[3] = 1.0f;
                finished = true;
            }
            t += a.step;
            finished = finished || (t > tend);
        }
    }
    if (inside) a.out[pix] = make_float4(out[0], out[1], out[2], out[3]);
    wave_add(a.counters + 0, n_rays);
    wave_add(a.counters + 1, n_samples);
    wave_add(a.counters + 2, n_samples);
    wave_add(a.counters + 3, n_shaded);
}

void launch_march_modes(const MarchArgs& a, int dtype, int filter, bool phong, hipStream_t st) {
    const dim3 grid((a.w + 15) / 16, (a.h + 15) / 16);
    if (dtype == 0) {
        if (filter == 0) {
            if (phong) hipLaunchKernelGGL((march_modes_kernel<uint8_t, 0, true>), grid, dim3(256), 0, st, a);
            else hipLaunchKernelGGL((march_modes_kernel<uint8_t, 0, false>), grid, dim3(256), 0, st, a);
        } else {
            if (phong) hipLaunchKernelGGL((march_modes_kernel<uint8_t, 1, true>), grid, dim3(256), 0, st, a);
            else hipLaunchKernelGGL((march_modes_kernel<uint8_t, 1, false>), grid, dim3(256), 0, st, a);
        }
    } else {
        if (filter == 0) {
            if (phong) hipLaunchKernelGGL((march_modes_kernel<uint16_t, 0, true>), grid, dim3(256), 0, st, a);
            else hipLaunchKernelGGL((march_modes_kernel<uint16_t, 0, false>), grid, dim3(256), 0, st, a);
        } else {
            if (phong) hipLaunchKernelGGL((march_modes_kernel<uint16_t, 1, true>), grid, dim3(256), 0, st, a);
            else hipLaunchKernelGGL((march_modes_kernel<uint16_t, 1, false>), grid, dim3(256), 0, st, a);
        }
    }
}

template <typename T, int FILTER, int CLASSIF, bool PHONG>
static void launch_t(const MarchArgs& a, bool count, hipStream_t st) {
    dim3 grid((a.w + 15) / 16, (a.h + 15) / 16);
    if (count)
        hipLaunchKernelGGL((march_ref_kernel<T, FILTER, CLASSIF, PHONG, true>), grid, dim3(256), 0,
                           st, a);
    else
        hipLaunchKernelGGL((march_ref_kernel<T, FILTER, CLASSIF, PHONG, false>), grid, dim3(256), 0,
                           st, a);
}

template <typename T>
static void launch_dtype(const MarchArgs& a, int filter, int classif, bool phong, bool count,
                         hipStream_t st) {
    if (classif == 2) {
        if (filter == 0) {
            if (phong) launch_t<T, 0, 2, true>(a, count, st); else launch_t<T, 0, 2, false>(a, count, st);
        } else {
            if (phong) launch_t<T, 1, 2, true>(a, count, st); else launch_t<T, 1, 2, false>(a, count, st);
        }
        return;
    }
    if (filter == 0) {
        if (classif == 0) {
            if (phong) launch_t<T, 0, 0, true>(a, count, st); else launch_t<T, 0, 0, false>(a, count, st);
        } else {
            if (phong) launch_t<T, 0, 1, true>(a, count, st); else launch_t<T, 0, 1, false>(a, count, st);
        }
    } else {
        if (classif == 0) {
            if (phong) launch_t<T, 1, 0, true>(a, count, st); else launch_t<T, 1, 0, false>(a, count, st);
        } else {
            if (phong) launch_t<T, 1, 1, true>(a, count, st); else launch_t<T, 1, 1, false>(a, count, st);
        }
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_march_ref(const MarchArgs& a, int dtype, int filter, int classif, bool phong,
                      bool count, hipStream_t st) {
    if (dtype == 0) launch_dtype<uint8_t>(a, filter, classif, phong, count, st);
    else launch_dtype<uint16_t>(a, filter, classif, phong, count, st);
}

}  // namespace hipdvr
