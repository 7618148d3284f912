// march_common.h -- per-sample arithmetic of the ray march, shared by the
// reference-layout and bricked kernels.
//
// Every float expression follows the evaluation order of Voreen's
// CPURaycaster (modules/base/processors/render/cpuraycaster.cpp:210-332) and
// VolumeRAM::getVoxelNormalizedLinear (src/core/datastructures/volume/
// volumeram.cpp:78-97).  The translation units that include this header are
// compiled with -ffp-contract=off, so no multiply-add is fused unless written
// as an explicit fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hipdvr {

constexpr float kErtOpacity = 0.95f;        // mod_raysetup.frag:36, cpuraycaster.cpp:310
constexpr int kMaxTfWidth = 4096;           // transfunc1dkeysproperty.cpp:95
constexpr int kDefaultLoopCap = 255 * 255;  // cpuraycaster.cpp:239
constexpr int kNumCounters = 6;
// shading models (values of hipdvr_params::shading)
constexpr int kShadeNone = 0, kShadePhong = 1, kShadePhongD = 2, kShadePhongS = 3, kShadePhongDA = 4,
              kShadePhongDS = 5, kShadeToon = 6, kShadeCookTorrance = 7, kShadeOrenNayar = 8, kShadeWard = 9;
// compositing modes (values of hipdvr_params::compositing)
constexpr int kCompDVR = 0, kCompMIP = 1, kCompMOP = 2, kCompMIDA = 3, kCompISO = 4, kCompFHP = 5,
              kCompFHN = 6;
// gradient modes (VolumeRaycaster "gradient" property, volumeraycaster.cpp:126-131,197-207)
constexpr int kGradCentral = 0, kGradForward = 1, kGradNone = 2, kGradFiltered = 3, kGradSobel = 4;

// Arguments of one march launch (plain data; passed by value as kernarg).
struct MarchArgs {
    const void* vox;            // linear voxels (reference kernel)
    const void* bricks;         // 4^3-block tiled voxels (bricked kernel)
    const void* cells;          // corner-replicated cell records (cell kernel)
    const uint8_t* brick_occ;   // per-brick occupancy (1 = may hold alpha > 0) for this TF
    const uint32_t* occ_bits;   // the same occupancy, one bit per brick (bit b&31 of word b>>5)
    uint32_t occ_words;
    uint32_t dims[3];
    uint32_t nbricks[3];
    uint32_t brick_log2;
    float fdims[3];             // (float)dims
    float fdm1[3];              // (float)dims - 1
    float rwm_scale, rwm_offset;
    float dom_lo, dom_hi;
    int32_t dom_unit;           // dom_hi - dom_lo == 1: the division is exact, x = rw - lo
    int32_t dtype;              // 0 uint8, 1 uint16
    int32_t tf_identity;        // rwm (1, 0) and TF domain [0, 1]: x = min(intensity, 1)
    int32_t shade_model;        // kShade*
    int32_t tf_width;
    const uint32_t* tf;         // RGBA8 packed little-endian (r | g<<8 | b<<16 | a<<24)
    const float4* preint;       // N*N, [fe*N + fs]
    int32_t preint_n;
    // 2D TF classification: gradient volume (same dims, gvol_channels interleaved
    // channels of u8 / u16 / f32 per voxel) and the opacity-corrected table
    const void* gvol;
    int32_t gvol_dtype, gvol_channels;
    const float4* tf2d;         // [y*tf2d_w + x]
    int32_t tf2d_w, tf2d_h;
    float tf2d_dom[4];          // intensity lo, hi; gradient magnitude lo, hi
    float step;
    int32_t loop_cap;
    // Phong
    float spacing[3], offset[3];
    float grad_off[3];          // 1/dims (datasetDimensionsRCP_)
    float grad_scale[3];        // (1/spacing) * 0.5
    float grad_raw[3];          // grad_scale * rwm_scale / (255 or 65535): differences of stored values
    float spacing_rcp[3];       // 1/spacing (forward differences)
    int32_t gradient;           // kGrad*
    float light_pos[3], light_amb[3], light_dif[3], light_spe[3], shininess;
    float cam_pos[3];
    // images
    const float4* entry;
    const float4* exit;
    float4* out;
    int32_t w, h;               // image size (entry/exit indexing)
    // launch grid in pixels: grid pixel (gx, gy) marches image pixel
    // (gx*stride + pix_offset, gy*stride + pix_offset), writes out[gy*grid_w + gx]
    // (the judge's re-march places grid pixels with judge_pixel and pix_block)
    int32_t grid_w, grid_h, stride, pix_offset, pix_block;
    const int* cond_flag;       // if set, the launch does nothing unless *cond_flag != 0
    int32_t compositing;        // kComp*
    float gamma, iso_value;     // MIDA gamma, ISO value
    unsigned long long* counters;  // [rays, samples, fetched, shaded, lane_slots, shade_slots]
    float alpha_corr[256];      // 1 - pow(1 - a/255, step*200) per alpha byte
};

// value / 255.f or value / 65535.f correctly rounded (the reference's
// getTypeAsFloat), as fma(fv, r_hi, fv * r_lo) with r_hi + r_lo = 1/range to
// about 48 bits: the quotient of an integer by the range never lies within
// that error of a rounding midpoint, so the single rounding of the fma gives
// the correctly rounded quotient (checked for every stored value).
template <typename T> struct VoxNorm;
template <> struct VoxNorm<uint8_t> {
    static __device__ __forceinline__ float apply(uint32_t v) {
        const float fv = (float)v;
        return __builtin_fmaf(fv, 0x1.010102p-8f, fv * -0x1.fdfdfep-33f);
    }
};
template <> struct VoxNorm<uint16_t> {
    static __device__ __forceinline__ float apply(uint32_t v) {
        const float fv = (float)v;
        return __builtin_fmaf(fv, 0x1.0001p-16f, fv * 0x1.0001p-48f);
    }
};

__device__ __forceinline__ float smaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sminf(float a, float b) { return (b < a) ? b : a; }
// tgt::iround (ext/tgt/tgt_math.h:213-225)
__device__ __forceinline__ int iround_tgt(float f) { return (int)floorf(f + 0.5f); }

// Texture -> clamped voxel coordinate (cpuraycaster.cpp:243 with
// VolumeBase::getTextureToVoxelMatrix, volumebase.cpp:221-224).
__device__ __forceinline__ float tex_to_voxel(float s, float fdim, float fdm1) {
    float v = (fdim * s) + (-0.5f);
    return sminf(smaxf(v, 0.0f), fdm1);
}

// Trilinear weights and corner coordinates (volumeram.cpp:79-86).
struct Corners {
    uint32_t lo[3];
    uint32_t d[3];      // hi - lo, 0 or 1
    float p[3];
};

__device__ __forceinline__ void corners(const float vp[3], const uint32_t dims[3], Corners& c) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float pa = smaxf(vp[i], 0.0f);
        c.p[i] = pa - floorf(pa);
        uint32_t lo = (uint32_t)pa;
        uint32_t hi = (uint32_t)ceilf(pa);
        uint32_t m = dims[i] - 1;
        hi = hi < m ? hi : m;
        lo = lo < m ? lo : m;
        c.lo[i] = lo;
        c.d[i] = hi - lo;
    }
}

// The 8-term sum in the reference order llB, lrB, urB, ulB, llF, lrF, urF, ulF.
__device__ __forceinline__ float trilerp(const float v[8], const float p[3]) {
    float ix = 1.f - p[0], iy = 1.f - p[1], iz = 1.f - p[2];
    float s = v[0] * ix * iy * iz;
    s = s + v[1] * p[0] * iy * iz;
    s = s + v[2] * p[0] * p[1] * iz;
    s = s + v[3] * ix * p[1] * iz;
    s = s + v[4] * ix * iy * p[2];
    s = s + v[5] * p[0] * iy * p[2];
    s = s + v[6] * p[0] * p[1] * p[2];
    s = s + v[7] * ix * p[1] * p[2];
    return s;
}

// TransFuncBase::realWorldToNormalized (transfuncbase.cpp:244-259).
__device__ __forceinline__ float rw_to_norm(float rw, float lo, float hi) {
    if (rw < lo) return 0.f;
    if (rw > hi) return 1.f;
    if (lo == hi) return 0.f;
    return (rw - lo) / (hi - lo);
}

// rw_to_norm for a domain of width exactly 1: (rw - lo) / 1 == rw - lo.
__device__ __forceinline__ float rw_to_norm_unit(float rw, float lo, float hi) {
    if (rw < lo) return 0.f;
    if (rw > hi) return 1.f;
    return rw - lo;
}

// Intensity -> TF table index (cpuraycaster.cpp:254,266 and :328-332).
__device__ __forceinline__ int tf_index(float inten, const MarchArgs& a) {
    if (a.tf_identity) {
        // rw = inten*1 + 0 == inten; intensities are >= 0, so the mapping
        // reduces to the upper clamp, and x*(W-1) needs no index clamp
        const float x = inten > 1.0f ? 1.0f : inten;
        return iround_tgt(x * (float)(a.tf_width - 1));
    }
    float rw = (inten * a.rwm_scale) + a.rwm_offset;
    float x = a.dom_unit ? rw_to_norm_unit(rw, a.dom_lo, a.dom_hi) : rw_to_norm(rw, a.dom_lo, a.dom_hi);
    int w1 = a.tf_width - 1;
    int idx = iround_tgt(x * (float)w1);
    return idx < 0 ? 0 : (idx > w1 ? w1 : idx);
}

__device__ __forceinline__ float norm_rw(float inten, const MarchArgs& a) {
    float rw = (inten * a.rwm_scale) + a.rwm_offset;
    return rw_to_norm(rw, a.dom_lo, a.dom_hi);
}

__device__ __forceinline__ int preint_index(float x, int n) {
    int i = iround_tgt(x * (float)(n - 1));
    return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// byte / 255.f, correctly rounded (same residual step as VoxNorm<uint8_t>).
__device__ __forceinline__ float byte_norm(uint32_t b) { return VoxNorm<uint8_t>::apply(b); }

// Phong shading, mod_shading.frag:134-151 with ka = kd = TF colour, ks = 1;
// a zero gradient gives N = 0.
__device__ __forceinline__ void phong(const MarchArgs& a, const float g[3], const float pos[3],
                                      float c[3]) {
    float gl = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    float N[3] = {0.f, 0.f, 0.f};
    if (gl > 0.f) {
        float r = 1.f / sqrtf(gl);
#pragma unroll
        for (int i = 0; i < 3; ++i) N[i] = g[i] * r;
    }
    float L[3], V[3], H[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { L[i] = a.light_pos[i] - pos[i]; V[i] = a.cam_pos[i] - pos[i]; }
    float vl = sqrtf(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    float rv = 1.f / vl;
#pragma unroll
    for (int i = 0; i < 3; ++i) V[i] *= rv;
    float d = sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) L[i] /= d;
#pragma unroll
    for (int i = 0; i < 3; ++i) H[i] = V[i] + L[i];
    float hl = sqrtf(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    float rh = 1.f / hl;
#pragma unroll
    for (int i = 0; i < 3; ++i) H[i] *= rh;
    float ndl = fmaxf(N[0] * L[0] + N[1] * L[1] + N[2] * L[2], 0.f);
    float ndh = fmaxf(N[0] * H[0] + N[1] * H[1] + N[2] * H[2], 0.f);
    float sp = powf(ndh, a.shininess);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float amb = c[i] * a.light_amb[i];
        float dif = (c[i] * a.light_dif[i]) * ndl;
        float spe = (1.0f * a.light_spe[i]) * sp;
        c[i] = ((0.0f + amb) + dif) + spe;
    }
}

// GLSL normalize() with a zero vector mapped to zero (undefined in GLSL).
__device__ __forceinline__ void normalize3(float v[3]) {
    const float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (l2 > 0.f) {
        const float r = 1.f / sqrtf(l2);
        v[0] *= r;
        v[1] *= r;
        v[2] *= r;
    } else {
        v[0] = v[1] = v[2] = 0.f;
    }
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// The shading models of mod_shading.frag as rc_singlevolume.frag:229 calls
// them through APPLY_SHADING (glsl.cpp:411-434): ka = kd = TF colour, ks = 1.
//   2 phongShadingD  :245-262   3 phongShadingS  :188-210   4 phongShadingDA :222-239
//   5 phongShadingDS :160-177   6 toonShading    :273-288 (3 shades)
//   7 cookTorranceShading :303-361   8 orenNayarShading :374-405   9 wardShading :419-483
// Phong itself (1) is phong() above.  acos arguments are clamped to [-1, 1]
// and a zero vector normalizes to zero, where GLSL leaves the result undefined.
__device__ __forceinline__ void shade_model(const MarchArgs& a, const float g[3], const float pos[3], float c[3]) {
    const int m = a.shade_model;
    if (m == kShadePhong) {
        phong(a, g, pos, c);
        return;
    }
    float N[3] = {g[0], g[1], g[2]};
    normalize3(N);
    float L[3], V[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L[i] = a.light_pos[i] - pos[i];
        V[i] = a.cam_pos[i] - pos[i];
    }
    float out[3] = {0.f, 0.f, 0.f};
    if (m == kShadeToon) {
        // L is not normalized here (mod_shading.frag:276)
        float ndl = fmaxf(dot3(N, L), 0.f);
        for (int i = 1; i <= 3; ++i) {
            if (ndl <= (float)i / 3.0f) {
                ndl = (float)i / 3.0f;
                break;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) out[i] = (c[i] * a.light_dif[i]) * ndl;
    } else if (m <= kShadePhongDS) {
        normalize3(V);
        const float d = sqrtf(dot3(L, L));
#pragma unroll
        for (int i = 0; i < 3; ++i) L[i] /= d;
        const float ndl = fmaxf(dot3(N, L), 0.f);
        float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
        normalize3(H);
        const float sp = powf(fmaxf(dot3(N, H), 0.f), a.shininess);
        const bool amb = m == kShadePhongDA;
        const bool dif = m == kShadePhongD || m == kShadePhongDA || m == kShadePhongDS;
        const bool spe = m == kShadePhongS || m == kShadePhongDS;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float s = 0.0f;
            if (amb) s = s + c[i] * a.light_amb[i];
            if (dif) s = s + (c[i] * a.light_dif[i]) * ndl;
            if (spe) s = s + (1.0f * a.light_spe[i]) * sp;
            out[i] = s;
        }
    } else {
        normalize3(L);
        normalize3(V);
        float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
        normalize3(H);
        if (m == kShadeCookTorrance) {
            const float rough = 1.0f - (a.shininess / 128.0f);
            const float s = a.shininess / 128.0f, dd = 1.0f - s;
            const float dw = sqrtf(dot3(L, L)) * 1.0f;
            const float NH = dot3(N, H), NV = dot3(N, V), NL = fmaxf(0.0000001f, dot3(N, L)), VH = dot3(V, H);
            const float F = 1.0f + (1.0f - 1.0f) * powf(1.0f - NV, 5.0f);
            const float G1 = (2.0f * NH * NL) / VH, G2 = (2.0f * NH * NV) / VH;
            const float G = fminf(1.0f, fmaxf(0.0f, fminf(G1, G2)));
            const float mq = rough * rough + 1e-5f, NHq = NH * NH + 1e-5f;
            const float D1 = 1.0f / (4.0f * mq * NHq * NHq);
            const float D2 = expf(-(1.0f - NHq) / (mq * NHq));
            const float D = D1 * D2;
            const float k = ((F * D) * G) / ((3.1415926f * NV) * NL);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float Rd = ((c[i] * a.light_dif[i]) * 2.0f) * NL;
                const float Rs = (k * 1.0f) * a.light_spe[i];
                out[i] = (0.0f + c[i] * a.light_amb[i]) + ((NL * dw) * (s * Rs + dd * Rd));
            }
        } else if (m == kShadeOrenNayar) {
            const float rough = a.shininess / 128.0f;
            const float VdotN = dot3(V, N), LdotN = dot3(L, N), NdotL = fmaxf(0.0000001f, dot3(N, L));
            float u[3], w[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                u[i] = V[i] - N[i] * VdotN;
                w[i] = L[i] - N[i] * LdotN;
            }
            const float cosPhi = dot3(u, w);
            const float av = acosf(fminf(fmaxf(VdotN, -1.0f), 1.0f)), al = acosf(fminf(fmaxf(LdotN, -1.0f), 1.0f));
            const float alpha = fmaxf(av, al), beta = fminf(av, al);
            const float mq = rough * rough;
            const float A = 1.0f - 0.5f * (mq / (mq + 0.57f));
            const float B = 0.45f * (mq / (mq + 0.09f));
            const float r = 0.0f + (A + ((B * fmaxf(0.0f, cosPhi)) * sinf(alpha)) * tanf(beta));
#pragma unroll
            for (int i = 0; i < 3; ++i) out[i] = (0.0f + c[i] * a.light_amb[i]) + ((c[i] * a.light_dif[i]) * NdotL) * r;
        } else {   // Ward
            float ax = a.shininess / 128.0f;
            float ay = ax * 2.0f;
            const float eps[3] = {1.0f, 0.0f, 0.0f};
            float X[3] = {N[1] * eps[2] - N[2] * eps[1], N[2] * eps[0] - N[0] * eps[2], N[0] * eps[1] - N[1] * eps[0]};
            normalize3(X);
            float Y[3] = {N[1] * X[2] - N[2] * X[1], N[2] * X[0] - N[0] * X[2], N[0] * X[1] - N[1] * X[0]};
            normalize3(Y);
            const float VdotN = dot3(V, N), LdotN = dot3(L, N), HdotN = dot3(H, N);
            const float HdotX = dot3(H, X), HdotY = dot3(H, Y), NdotL = fmaxf(0.0000001f, dot3(N, L));
            ax += 0.0000001f;
            ay += 0.0000001f;
            float A = HdotX / ax;
            A *= A;
            float B = HdotY / ay;
            B *= B;
            const float expo = -2.0f * ((A + B) / (1.0f + HdotN));
            float nenner = 4.0f * 3.14159f;
            nenner *= ax;
            nenner *= ay;
            nenner *= sqrtf(LdotN * VdotN);
            // a NaN (negative radicand) takes the max branch, which keeps 1e-7
            if (nenner < 0.0f) nenner = fminf(0.0000001f, nenner);
            else nenner = fmaxf(0.0000001f, nenner);
            const float e = expf(expo) / nenner;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float Rd = c[i] * a.light_dif[i];
                const float Rs = (1.0f * a.light_spe[i]) * e;
                out[i] = (0.0f + c[i] * a.light_amb[i]) + NdotL * (Rs + Rd);
            }
        }
    }
    c[0] = out[0];
    c[1] = out[1];
    c[2] = out[2];
}

// The same Phong terms with hardware reciprocal square root and exp2/log2
// (about 1 ulp each) and fused multiply-adds: for kernels whose shading is
// checked against the FP64 oracle by tolerance.  A zero gradient gives N = 0; N.H = 0 gives a
// zero specular term.
__device__ __forceinline__ void phong_fast(const MarchArgs& a, const float g[3], const float pos[3],
                                           float c[3]) {
    const float gl = __builtin_fmaf(g[2], g[2], __builtin_fmaf(g[1], g[1], g[0] * g[0]));
    const float rn = gl > 0.f ? __builtin_amdgcn_rsqf(gl) : 0.f;
    float L[3], V[3], H[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { L[i] = a.light_pos[i] - pos[i]; V[i] = a.cam_pos[i] - pos[i]; }
    const float rv = __builtin_amdgcn_rsqf(__builtin_fmaf(V[2], V[2], __builtin_fmaf(V[1], V[1], V[0] * V[0])));
    const float rl = __builtin_amdgcn_rsqf(__builtin_fmaf(L[2], L[2], __builtin_fmaf(L[1], L[1], L[0] * L[0])));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L[i] *= rl;
        H[i] = __builtin_fmaf(V[i], rv, L[i]);   // normalize(V) + normalize(L)
    }
    const float rh = __builtin_amdgcn_rsqf(__builtin_fmaf(H[2], H[2], __builtin_fmaf(H[1], H[1], H[0] * H[0])));
    const float ndl = fmaxf(__builtin_fmaf(g[2], L[2], __builtin_fmaf(g[1], L[1], g[0] * L[0])) * rn, 0.f);
    const float ndh = fmaxf(__builtin_fmaf(g[2], H[2], __builtin_fmaf(g[1], H[1], g[0] * H[0])) * (rn * rh), 0.f);
    const float sp = ndh > 0.f ? __builtin_amdgcn_exp2f(a.shininess * __builtin_amdgcn_logf(ndh)) : 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)   // ka = kd = c: c * (amb + dif * N.L) + spe * sp
        c[i] = __builtin_fmaf(c[i], __builtin_fmaf(a.light_dif[i], ndl, a.light_amb[i]), a.light_spe[i] * sp);
}

// fixClipBorderGradient, mod_gradients.frag:39-46 (neighbours clamp to edge).
__device__ __forceinline__ void clip_border_gradient(const MarchArgs& a, int x, int y,
                                                     const float s[3], const float dir[3],
                                                     float g[3]) {
    int x2 = x + 2 < a.w ? x + 2 : a.w - 1;
    int y2 = y + 2 < a.h ? y + 2 : a.h - 1;
    float4 e0 = a.entry[(size_t)y * a.w + x2];
    float4 e1 = a.entry[(size_t)y2 * a.w + x];
    float v0[3] = {e0.x - s[0], e0.y - s[1], e0.z - s[2]};
    float v1[3] = {e1.x - s[0], e1.y - s[1], e1.z - s[2]};
    float l0 = sqrtf(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    float l1 = sqrtf(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    if (l0 > 0.f) { float r = 1.f / l0; v0[0] *= r; v0[1] *= r; v0[2] *= r; }
    if (l1 > 0.f) { float r = 1.f / l1; v1[0] *= r; v1[1] *= r; v1[2] *= r; }
    g[0] = v0[1] * v1[2] - v0[2] * v1[1];
    g[1] = v0[2] * v1[0] - v0[0] * v1[2];
    g[2] = v0[0] * v1[1] - v0[1] * v1[0];
    float dd = g[0] * dir[0] + g[1] * dir[1] + g[2] * dir[2];
    if (!(dd < 0.f)) { g[0] = -g[0]; g[1] = -g[1]; g[2] = -g[2]; }
}

// Ray setup, cpuraycaster.cpp:218-231.
__device__ __forceinline__ float ray_setup(const float first[3], const float last[3], float step,
                                           float dir[3]) {
    dir[0] = last[0] - first[0];
    dir[1] = last[1] - first[1];
    dir[2] = last[2] - first[2];
    bool dz = (dir[0] == 0.f && dir[1] == 0.f && dir[2] == 0.f);
    bool lnz = !(last[0] == 0.f && last[1] == 0.f && last[2] == 0.f);
    bool fnz = !(first[0] == 0.f && first[1] == 0.f && first[2] == 0.f);
    if (dz && lnz && fnz) return step / 2.0f;
    float lsq = 0.0f;
    lsq += dir[0] * dir[0];
    lsq += dir[1] * dir[1];
    lsq += dir[2] * dir[2];
    float len = sqrtf(lsq);
    float rl = 1.0f / len;
    dir[0] *= rl;
    dir[1] *= rl;
    dir[2] *= rl;
    return len;
}

__device__ __forceinline__ bool is_background(float4 f, float4 l) {
    return f.x == 0.f && f.y == 0.f && f.z == 0.f && f.w == 0.f && l.x == 0.f && l.y == 0.f &&
           l.z == 0.f && l.w == 0.f;
}

// Ray parameter below which every sample lies in the (empty) brick holding
// the voxel position vp: the smallest exit parameter over the three axes of
// the brick's voxel slab shrunk by 0.05 voxel (far above the float error of
// the sample position), in unclamped voxel space where the first and last
// brick of an axis extend to infinity (positions are clamped to the volume).
__device__ __forceinline__ float empty_brick_limit(const MarchArgs& a, const float first[3],
                                                   const float dir[3], const float vp[3],
                                                   uint32_t bl) {
    float tl = INFINITY;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const uint32_t b = (uint32_t)vp[i] >> bl;
        const float B = (float)(1u << bl);
        if (dir[i] > 0.0f && b + 1 < a.nbricks[i]) {
            const float vb = (float)(b + 1) * B - 0.05f;           // voxel coordinate bound
            const float sb = (vb + 0.5f) / a.fdims[i];               // texture coordinate
            tl = fminf(tl, (sb - first[i]) / dir[i]);
        } else if (dir[i] < 0.0f && b > 0) {
            const float vb = (float)b * B + 0.05f;
            const float sb = (vb + 0.5f) / a.fdims[i];
            tl = fminf(tl, (sb - first[i]) / dir[i]);
        }
    }
    return tl;
}

// empty_brick_limit from the reciprocals rdir = 1/dir (0: the axis bounds
// nothing) and rfd = 1/fdims, each within a few ulp: the limit moves by far
// less than the 0.05-voxel margin.
__device__ __forceinline__ float empty_brick_limit_rcp(const MarchArgs& a, const float first[3],
                                                       const float rdir[3], const float rfd[3],
                                                       const float vp[3], uint32_t bl) {
    float tl = INFINITY;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const uint32_t b = (uint32_t)vp[i] >> bl;
        const float B = (float)(1u << bl);
        if (rdir[i] > 0.0f && b + 1 < a.nbricks[i]) {
            const float vb = (float)(b + 1) * B - 0.05f;
            tl = fminf(tl, ((vb + 0.5f) * rfd[i] - first[i]) * rdir[i]);
        } else if (rdir[i] < 0.0f && b > 0) {
            const float vb = (float)b * B + 0.05f;
            tl = fminf(tl, ((vb + 0.5f) * rfd[i] - first[i]) * rdir[i]);
        }
    }
    return tl;
}

// 64-lane maximum of a per-lane counter, times 64, added once per wavefront.
__device__ __forceinline__ void wave_add_max64(unsigned long long* dst, unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, 64ull * v);
}

// 64-lane sum of a per-lane counter; lane 0 adds it to the global counter.
__device__ __forceinline__ void wave_add(unsigned long long* dst, unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// Image column (or row) of the judge's grid column g: blocks of `block`
// consecutive pixels, one block in every stride * block pixels, from offs
// (block 1: every stride-th pixel).
__host__ __device__ inline int judge_pixel(int g, int stride, int block, int offs) {
    return (g / block) * (stride * block) + offs + g % block;
}

// Image tiles (16x16 pixels, one workgroup each) in XCD-compact order.  The
// hardware hands consecutive workgroups to the 8 XCDs in turn (workgroup b
// to XCD b % 8).  Here the image is cut into chip regions of 4 x 2 XCD
// regions, each XCD region kXcdRegionW x kXcdRegionH tiles; the i-th
// workgroup of XCD x marches tile i of XCD region x of chip region
// i / (W*H), chip regions from the middle of the image out (center_out).
// So the rays one XCD has in flight lie in one compact region and share its
// L2, and the 8 XCDs march neighbouring regions, so the rays in flight on
// the chip cover one compact image area and share the Infinity Cache.  Measured on 512^3 / 1024^2 with
// Phong: plain dispatch order 1.90 ms; each XCD on 2x2-tile blocks taken in
// turn 1.79 ms; XCD regions of 4x4 tiles 1.74 ms, 8x4 1.73 ms, 8x8 1.79
// ms, 16x8 (chip regions as wide as the image) 2.74 ms; one contiguous
// image band per XCD 2.74 ms.  On 128^3 / 256^2 (16 x 16 tiles) the 2x2
// blocks keep all XCDs busy (51,220 frames/s against 46,139 with regions).
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdRegionW = 8, kXcdRegionH = 4;
constexpr uint32_t kChipRegionW = 4 * kXcdRegionW, kChipRegionH = 2 * kXcdRegionH;
// Images smaller than one chip region would leave XCDs without tiles; there
// each XCD takes 2x2-tile blocks in turn instead (block b % 8 of every
// round of 8 blocks, blocks row-major).
constexpr uint32_t kXcdBlock = 2;

__host__ __device__ inline bool xcd_regions(uint32_t tiles_x, uint32_t tiles_y) {
    return tiles_x >= kChipRegionW && tiles_y >= kChipRegionH;
}

// Workgroups to launch for tiles_x x tiles_y tiles (whole chip regions or
// whole rounds of blocks; the workgroups of tiles outside the image exit at
// once).
__host__ __device__ inline uint32_t xcd_tile_grid(uint32_t tiles_x, uint32_t tiles_y) {
    if (xcd_regions(tiles_x, tiles_y)) {
        const uint32_t cx = (tiles_x + kChipRegionW - 1) / kChipRegionW;
        const uint32_t cy = (tiles_y + kChipRegionH - 1) / kChipRegionH;
        return kXcds * kXcdRegionW * kXcdRegionH * cx * cy;
    }
    const uint32_t bx = (tiles_x + kXcdBlock - 1) / kXcdBlock, by = (tiles_y + kXcdBlock - 1) / kXcdBlock;
    return kXcds * kXcdBlock * kXcdBlock * ((bx * by + kXcds - 1) / kXcds);
}

// The k-th of n rows (or columns) of chip regions, from the middle out:
// c, c + 1, c - 1, c + 2, ... (c = (n - 1) / 2), then the rest of the longer
// side.  Long rays through the middle of the volume start first and the
// cheap edges fill the end of the launch (1.74 -> 1.71 ms on 512^3 / 1024^2
// with Phong), while the regions in flight stay next to each other.
__device__ __forceinline__ uint32_t center_out(uint32_t k, uint32_t n) {
    const uint32_t c = (n - 1) / 2;   // c rows below the middle one, n - 1 - c >= c above
    if (k == 0) return c;
    if (k <= 2 * c) return (k & 1) ? c + (k + 1) / 2 : c - (k + 1) / 2;
    return c + (k - c);
}

// Tile (tx, ty) of workgroup b; false for a workgroup with no tile.
__device__ __forceinline__ bool xcd_tile(uint32_t b, uint32_t tiles_x, uint32_t tiles_y, uint32_t& tx,
                                         uint32_t& ty) {
    const uint32_t xcd = b % kXcds, slot = b / kXcds;
    if (xcd_regions(tiles_x, tiles_y)) {
        constexpr uint32_t R = kXcdRegionW * kXcdRegionH;
        const uint32_t chip = slot / R, j = slot % R;
        const uint32_t cx_n = (tiles_x + kChipRegionW - 1) / kChipRegionW;
        const uint32_t cy_n = (tiles_y + kChipRegionH - 1) / kChipRegionH;
        const uint32_t cx = center_out(chip % cx_n, cx_n), cy = center_out(chip / cx_n, cy_n);
        tx = cx * kChipRegionW + (xcd % 4) * kXcdRegionW + j % kXcdRegionW;
        ty = cy * kChipRegionH + (xcd / 4) * kXcdRegionH + j / kXcdRegionW;
    } else {
        constexpr uint32_t S = kXcdBlock * kXcdBlock;
        const uint32_t block = (slot / S) * kXcds + xcd, j = slot % S;
        const uint32_t bx = (tiles_x + kXcdBlock - 1) / kXcdBlock;
        tx = (block % bx) * kXcdBlock + j % kXcdBlock;
        ty = (block / bx) * kXcdBlock + j / kXcdBlock;
    }
    return tx < tiles_x && ty < tiles_y;
}

}  // namespace hipdvr
