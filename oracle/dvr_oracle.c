/*
 * dvr_oracle.c -- CPU oracle for the DVR ray march (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, always as the checker.  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off -fopenmp), which keeps every float expression in the
 * reference's evaluation order with no fused multiply-add.
 *
 * How each function is pinned (tests/test_oracle.py unless noted):
 *   dvr_ray_setup,      bit for bit against tests/golden/tgt_ray_samples.txt:
 *   dvr_sample_voxel    tend, direction, the t recurrence, the texture-to-voxel
 *                       position, its clamp and rounding, as the reference's
 *                       own tgt computes them (oracle/tgt_probe.cpp, 96 rays).
 *   dvr_sample_linear   affine fields are reproduced exactly at any position;
 *                       integer positions return the voxel; clamp at faces;
 *                       brute-force FP64 trilinear on random 2^3..5^3 volumes.
 *   dvr_tf_build        the standard 2-key ramp (transfunc1dkeys.cpp:413-421)
 *                       gives entry i = trunc(255*i/(W-1)); constant TF; split
 *                       keys; threshold zeroing; alpha modes; the nucleon and
 *                       walnut .tfi fixtures shipped with the reference.
 *   dvr_cpu_render      constant volume: A_n = 1-(1-a')^n closed form and the
 *                       ERT index; ray-length sample count floor(tend/step)+1
 *                       (float accumulation); degenerate ray = 1 sample;
 *                       background pixels exactly 0; single opaque slab gives the
 *                       first-hit colour; FP64 oracle envelope (PSNR).
 *   dvr_fp64_render     same closed forms in double.
 *   dvr_preint_build    constant TF: diagonal = opacity-corrected TF; integral
 *                       variant = 1-exp(-a*200*step) for constant alpha.
 *   dvr_entry_exit      analytic: every hit point lies on the unit-cube surface
 *                       (or on the near plane), lies on the pixel's ray, and
 *                       misses are exactly zero; brute-force ray marching.
 *   Phong               single-term checks (ambient only, N parallel to L) and
 *                       the gradient of an affine field.
 *   compositing modes   MIP on a linear ramp along the ray = the colour of the
 *                       TF entry of the largest intensity; MOP = the entry of the
 *                       largest TF alpha; ISO on a ramp = the first sample whose
 *                       alpha is within 0.02 of the iso value; FHP = the entry
 *                       point of a ray entering an opaque volume (t = 0 sample);
 *                       FHN of an affine field = its normalized down-gradient
 *                       mapped to [0,1]; MIDA with gamma = 1 equals MIP with the
 *                       corrected alpha and with a constant volume the
 *                       closed-form recurrence (delta = voxel at every sample).
 *                       A zero FHN gradient gives (0.5, 0.5, 0.5) where GLSL's
 *                       normalize is undefined: parity unpinned there.
 *   shading models      N = L = V (and N = H) closed forms: each Phong term
 *                       alone, the cos(theta) diffuse fall-off, toon's i/3
 *                       quantization on the unnormalized N.L, Cook-Torrance
 *                       with F = G = 1 and Beckmann D at NH = 1, Oren-Nayar's A
 *                       term at alpha = beta = 0, Ward's 1/(4 pi ax ay) peak.
 *   gradient modes      affine fields give the same down-gradient in every
 *                       mode; x^2 (forward: -(2x+1)), x*y^2 (Sobel smooths the
 *                       y^2 across planes: +1/2), x^3 (filtered blend 1/4 :
 *                       3/4: -(3x^2+3.25)); spacing and rwm scaling; "none"
 *                       = (rw-0.5)*2 and (offset-0.5)*2.
 *   2D TF               single-sample rays: the texel at the truncated
 *                       (intensity, gradient-magnitude) position of affine
 *                       fields, opacity-corrected; f32/u8/u16 gradient volumes,
 *                       3 and 4 channels, both filters.
 */
#include "dvr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* cpuraycaster.cpp:42 and mod_compositing.frag:40 */
#define SAMPLING_BASE_INTERVAL_RCP 200.0f
/* mod_raysetup.frag:36; cpuraycaster.cpp:310 */
#define ERT_OPACITY 0.95f

/* ------------------------------------------------------------------------ */
/* PRNG                                                                      */
/* ------------------------------------------------------------------------ */
uint64_t dvr_splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double dvr_uniform01(uint64_t* s) {
    return (double)(dvr_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

/* tgt::iround, ext/tgt/tgt_math.h:213-225: (int)floor(f + 0.5f) */
static inline int iroundf_tgt(float f) { return (int)floorf(f + 0.5f); }
static inline int iround_d(double d) { return (int)floor(d + 0.5); }

/* std::max / std::min as used by tgt::clamp (vector.h:841-915) */
static inline float smaxf(float a, float b) { return (a < b) ? b : a; }
static inline float sminf(float a, float b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Transfer function                                                         */
/* ------------------------------------------------------------------------ */

/* TransFunc1DKeys::getMappingForValueUByte, transfunc1dkeys.cpp:119-154 */
static void tf_map_ubyte(const dvr_tf_key* keys, int n, float value, uint8_t out[4]) {
    if (n <= 0) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    value = (value < 0.f) ? 0.f : value;
    value = (value > 1.f) ? 1.f : value;
    int k = 0;
    while (k < n && value > keys[k].intensity) k++;
    if (k == 0) {
        memcpy(out, keys[0].color_l, 4);
    } else if (k == n) {
        memcpy(out, keys[n - 1].color_r, 4);
    } else {
        const dvr_tf_key* L = &keys[k - 1];
        const dvr_tf_key* R = &keys[k];
        float fraction = (value - L->intensity) / (R->intensity - L->intensity);
        for (int c = 0; c < 4; ++c) {
            /* GLubyte += int * float: float sum, truncated back to the byte */
            int diff = (int)R->color_l[c] - (int)L->color_r[c];
            float s = (float)L->color_r[c] + (float)diff * fraction;
            out[c] = (uint8_t)s;
        }
    }
}

/* TransFunc1D::applyGammaToIntensity, transfunc1d.cpp:144-149 */
static inline float apply_gamma(float v, float gamma) {
    return (gamma != 1.f) ? powf(v, gamma) : v;
}

/* TransFunc1D::createTextureData (transfunc1d.cpp:85-116) + applyAlpha
 * (transfuncbase.cpp:177-210) for the TF_UBYTE data type. */
void dvr_tf_build(const dvr_tf_key* keys, int32_t nkeys, int32_t width, float gamma,
                  float thr_lo, float thr_hi, int32_t alpha_mode, uint8_t* out) {
    int front_end = iroundf_tgt(thr_lo * (float)width);
    int back_start = iroundf_tgt(thr_hi * (float)width);
    for (int x = 0; x < width; ++x) {
        uint8_t* o = out + 4 * (size_t)x;
        if (x < front_end || x > back_start) {
            o[0] = o[1] = o[2] = o[3] = 0;
        } else {
            float v = (float)x / (float)(width - 1);
            tf_map_ubyte(keys, nkeys, apply_gamma(v, gamma), o);
        }
        if (alpha_mode == DVR_ALPHA_ONE) {
            if (o[3]) o[3] = 255;
        } else if (alpha_mode == DVR_ALPHA_ZERO) {
            o[3] = 0;
        }
    }
}

/* cpuraycaster.cpp:291: color.a = 1 - pow(1 - color.a, step * 200), float
 * arguments (the float overload of pow). */
float dvr_opacity_correct(float alpha, float step) {
    return 1.f - powf(1.f - alpha, step * SAMPLING_BASE_INTERVAL_RCP);
}

/* ------------------------------------------------------------------------ */
/* Pre-integration table, preintegrationtable.cpp:67-204                      */
/* ------------------------------------------------------------------------ */
void dvr_preint_build(const dvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo,
                      float thr_hi, int32_t alpha_mode, int32_t n, float step,
                      int32_t use_integral, float* table) {
    float* tfb = (float*)calloc((size_t)n * 4, sizeof(float));
    int front_end = iroundf_tgt(thr_lo * (float)n);
    int back_start = iroundf_tgt(thr_hi * (float)n);
    for (int i = front_end; i < back_start && i < n; ++i) {
        if (i < 0) continue;
        float intensity = (float)i / (float)(n - 1);
        uint8_t c8[4];
        tf_map_ubyte(keys, nkeys, apply_gamma(intensity, gamma), c8);
        float v[4];
        for (int c = 0; c < 4; ++c) v[c] = (float)c8[c] / 255.f;
        if (alpha_mode == DVR_ALPHA_ONE && v[3] != 0.f) v[3] = 1.f;
        else if (alpha_mode == DVR_ALPHA_ZERO) v[3] = 0.f;
        memcpy(tfb + 4 * i, v, sizeof v);
    }
    const float e_full = step * 200.0f;
    if (!use_integral) {
        #pragma omp parallel for schedule(dynamic)
        for (int sb = 0; sb < n; ++sb) {
            for (int sf = 0; sf < n; ++sf) {
                float* dst = table + 4 * ((size_t)sb * n + sf);
                if (sb != sf) {
                    float scale = 1.0f / (fabsf((float)(sb - sf)) + 1);
                    int incr = (sb < sf) ? -1 : 1;
                    float r = 0.f, g = 0.f, b = 0.f, a = 0.f;
                    /* preintegrationtable.cpp:115 compares the float alpha with the
                     * double 0.95 (a composite alpha of exactly 0.95f continues) */
                    for (int s = sf; (incr == 1 ? s <= sb : s >= sb) && ((double)a < 0.95); s += incr) {
                        const float* cc = tfb + 4 * s;
                        float ca = cc[3];
                        if (ca > 0.0f) {
                            ca = 1.f - powf(1.f - ca, step * 200.0f * scale);
                            float w = (1.0f - a) * ca;
                            float nr = r + w * cc[0];
                            float ng = g + w * cc[1];
                            float nb = b + w * cc[2];
                            a = a + (1.0f - a) * ca;
                            r = nr; g = ng; b = nb;
                        }
                    }
                    float den = smaxf(a, 0.001f);   /* std::max(result.a, 0.001f) */
                    r /= den; g /= den; b /= den;
                    dst[0] = sminf(smaxf(r, 0.f), 1.f);
                    dst[1] = sminf(smaxf(g, 0.f), 1.f);
                    dst[2] = sminf(smaxf(b, 0.f), 1.f);
                    dst[3] = sminf(smaxf(a, 0.f), 1.f);
                } else {
                    const float* cc = tfb + 4 * sf;
                    for (int c = 0; c < 4; ++c) dst[c] = sminf(smaxf(cc[c], 0.f), 1.f);
                    dst[3] = 1.f - powf(1.f - dst[3], e_full);
                }
            }
        }
    } else {
        float* intf = (float*)calloc((size_t)n * 4, sizeof(float));
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < n; ++i) {
            const float* cc = tfb + 4 * i;
            if (cc[3] > 0.0f) {
                acc[0] += cc[0] * cc[3];
                acc[1] += cc[1] * cc[3];
                acc[2] += cc[2] * cc[3];
                acc[3] += cc[3];
            }
            memcpy(intf + 4 * i, acc, sizeof acc);
        }
        #pragma omp parallel for schedule(static)
        for (int sb = 0; sb < n; ++sb) {
            for (int sf = 0; sf < n; ++sf) {
                int smin = sb < sf ? sb : sf;
                int smax = sb < sf ? sf : sb;
                float col[4];
                if (smax != smin) {
                    float factor = step * 200.f / (float)(smax - smin);
                    for (int c = 0; c < 3; ++c)
                        col[c] = (intf[4 * smax + c] - intf[4 * smin + c]) * factor;
                    col[3] = 1.f - expf(-(intf[4 * smax + 3] - intf[4 * smin + 3]) * factor);
                    float den = smaxf(col[3], 0.001f);
                    for (int c = 0; c < 3; ++c) col[c] /= den;
                } else {
                    memcpy(col, tfb + 4 * smin, sizeof col);
                    col[3] = 1.f - powf(1.f - col[3], e_full);
                }
                float* dst = table + 4 * ((size_t)sb * n + sf);
                for (int c = 0; c < 4; ++c) dst[c] = sminf(smaxf(col[c], 0.f), 1.f);
            }
        }
        free(intf);
    }
    free(tfb);
}

/* ------------------------------------------------------------------------ */
/* Entry/exit points for the unit-cube proxy                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    float o[3];     /* camera position in texture space */
    float f[3];     /* view direction, texture space */
    float r[3];     /* right * tan(fovy/2) * aspect, texture space */
    float u[3];     /* up * tan(fovy/2), texture space */
    float near_t;   /* near-plane ray parameter (dot(d_world, f_world) == 1) */
} raygen;

static void raygen_setup(const dvr_camera* cam, const float llf[3], const float urb[3],
                         int w, int h, raygen* g) {
    double f[3], up[3], r[3], u[3], n;
    for (int i = 0; i < 3; ++i) { f[i] = (double)cam->focus[i] - cam->eye[i]; up[i] = cam->up[i]; }
    n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int i = 0; i < 3; ++i) f[i] /= n;
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    n = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int i = 0; i < 3; ++i) r[i] /= n;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    /* tgt::Camera::getProjectionMatrix (camera.cpp:203-214), frustum ratio
     * 1: fovy clamped to [6, 175] degrees spans the image height when the
     * window is at least as wide as high, else its width */
    double fovy = cam->fovy_deg < 6.0f ? 6.0 : (cam->fovy_deg > 175.0f ? 175.0 : (double)cam->fovy_deg);
    double th = tan(fovy * 0.5 * 3.14159265358979323846 / 180.0);
    double aspect = (double)w / (double)h;
    double th_u = aspect >= 1.0 ? th : th / aspect;   /* half height on the unit view distance */
    double th_r = th_u * aspect;                      /* half width */
    for (int i = 0; i < 3; ++i) {
        double inv = 1.0 / ((double)urb[i] - (double)llf[i]);
        g->o[i] = (float)(((double)cam->eye[i] - (double)llf[i]) * inv);
        g->f[i] = (float)(f[i] * inv);
        g->r[i] = (float)(r[i] * th_r * inv);
        g->u[i] = (float)(u[i] * th_u * inv);
    }
    g->near_t = cam->near_dist;
}

static inline float clamp01(float v) { return sminf(smaxf(v, 0.f), 1.f); }

static void raygen_pixel(const raygen* g, int x, int y, int w, int h, int quantize,
                         float* en, float* ex) {
    float nx = (2.0f * ((float)x + 0.5f)) / (float)w - 1.0f;
    float ny = (2.0f * ((float)y + 0.5f)) / (float)h - 1.0f;
    float d[3];
    float tn = -INFINITY, tf = INFINITY;
    int miss = 0;
    for (int i = 0; i < 3; ++i) {
        d[i] = (g->f[i] + nx * g->r[i]) + ny * g->u[i];
        if (d[i] != 0.0f) {
            float inv = 1.0f / d[i];
            float ta = (0.0f - g->o[i]) * inv;
            float tb = (1.0f - g->o[i]) * inv;
            float lo = fminf(ta, tb), hi = fmaxf(ta, tb);
            tn = fmaxf(tn, lo);
            tf = fminf(tf, hi);
        } else if (g->o[i] < 0.0f || g->o[i] > 1.0f) {
            miss = 1;
        }
    }
    float ts = fmaxf(tn, g->near_t);
    if (miss || !(tf > ts)) {
        for (int c = 0; c < 4; ++c) { en[c] = 0.f; ex[c] = 0.f; }
        return;
    }
    for (int i = 0; i < 3; ++i) {
        float a = clamp01(g->o[i] + ts * d[i]);
        float b = clamp01(g->o[i] + tf * d[i]);
        if (quantize) {
            a = floorf(a * 65535.0f + 0.5f) / 65535.0f;
            b = floorf(b * 65535.0f + 0.5f) / 65535.0f;
        }
        en[i] = a;
        ex[i] = b;
    }
    en[3] = 1.0f;
    ex[3] = 1.0f;
}

void dvr_entry_exit(const dvr_camera* cam, const float llf[3], const float urb[3],
                    int32_t w, int32_t h, int32_t quantize16, float* entry, float* exit_) {
    raygen g;
    raygen_setup(cam, llf, urb, w, h, &g);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t p = (size_t)y * w + x;
            raygen_pixel(&g, x, y, w, h, quantize16, entry + 4 * p, exit_ + 4 * p);
        }
}

/* ------------------------------------------------------------------------ */
/* Volume access                                                             */
/* ------------------------------------------------------------------------ */

/* VolumeAtomic<T>::getVoxelNormalized (volumeatomic.h:531-536,599-608) with
 * getTypeAsFloat (volumeelement.h:1261-1276): value / rangeMax. */
static inline float voxel_norm(const dvr_volume* v, size_t x, size_t y, size_t z) {
    size_t idx = z * (size_t)v->dims[0] * (size_t)v->dims[1] + y * (size_t)v->dims[0] + x;
    if (v->dtype == DVR_U8)
        return (float)((const uint8_t*)v->data)[idx] / 255.0f;
    return (float)((const uint16_t*)v->data)[idx] / 65535.0f;
}

static inline double voxel_norm_d(const dvr_volume* v, size_t x, size_t y, size_t z) {
    size_t idx = z * (size_t)v->dims[0] * (size_t)v->dims[1] + y * (size_t)v->dims[0] + x;
    if (v->dtype == DVR_U8)
        return (double)((const uint8_t*)v->data)[idx] / 255.0;
    return (double)((const uint16_t*)v->data)[idx] / 65535.0;
}

/* VolumeRAM::getVoxelNormalizedLinear, volumeram.cpp:78-97.  The sum keeps
 * the reference order llB, lrB, urB, ulB, llF, lrF, urF, ulF and every term
 * is ((v * wx) * wy) * wz. */
float dvr_sample_linear(const dvr_volume* v, float px_, float py_, float pz_) {
    float pa[3] = { smaxf(px_, 0.0f), smaxf(py_, 0.0f), smaxf(pz_, 0.0f) };
    float p[3];
    size_t lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        p[i] = pa[i] - floorf(pa[i]);
        lo[i] = (size_t)pa[i];
        hi[i] = (size_t)ceilf(pa[i]);
        size_t m = (size_t)v->dims[i] - 1;
        hi[i] = hi[i] < m ? hi[i] : m;
        lo[i] = lo[i] < m ? lo[i] : m;
    }
    float ix = 1.f - p[0], iy = 1.f - p[1], iz = 1.f - p[2];
    return voxel_norm(v, lo[0], lo[1], lo[2]) * ix * iy * iz
         + voxel_norm(v, hi[0], lo[1], lo[2]) * p[0] * iy * iz
         + voxel_norm(v, hi[0], hi[1], lo[2]) * p[0] * p[1] * iz
         + voxel_norm(v, lo[0], hi[1], lo[2]) * ix * p[1] * iz
         + voxel_norm(v, lo[0], lo[1], hi[2]) * ix * iy * p[2]
         + voxel_norm(v, hi[0], lo[1], hi[2]) * p[0] * iy * p[2]
         + voxel_norm(v, hi[0], hi[1], hi[2]) * p[0] * p[1] * p[2]
         + voxel_norm(v, lo[0], hi[1], hi[2]) * ix * p[1] * p[2];
}

static double sample_linear_d(const dvr_volume* v, float px_, float py_, float pz_) {
    float pa[3] = { smaxf(px_, 0.0f), smaxf(py_, 0.0f), smaxf(pz_, 0.0f) };
    double p[3];
    size_t lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        p[i] = (double)pa[i] - floor((double)pa[i]);
        lo[i] = (size_t)pa[i];
        hi[i] = (size_t)ceilf(pa[i]);
        size_t m = (size_t)v->dims[i] - 1;
        hi[i] = hi[i] < m ? hi[i] : m;
        lo[i] = lo[i] < m ? lo[i] : m;
    }
    double ix = 1.0 - p[0], iy = 1.0 - p[1], iz = 1.0 - p[2];
    return voxel_norm_d(v, lo[0], lo[1], lo[2]) * ix * iy * iz
         + voxel_norm_d(v, hi[0], lo[1], lo[2]) * p[0] * iy * iz
         + voxel_norm_d(v, hi[0], hi[1], lo[2]) * p[0] * p[1] * iz
         + voxel_norm_d(v, lo[0], hi[1], lo[2]) * ix * p[1] * iz
         + voxel_norm_d(v, lo[0], lo[1], hi[2]) * ix * iy * p[2]
         + voxel_norm_d(v, hi[0], lo[1], hi[2]) * p[0] * iy * p[2]
         + voxel_norm_d(v, hi[0], hi[1], hi[2]) * p[0] * p[1] * p[2]
         + voxel_norm_d(v, lo[0], hi[1], hi[2]) * ix * p[1] * p[2];
}

/* Channel c of the gradient volume, normalized (getVoxelNormalized(pos,
 * channel), volumeatomic.h:531-536: float channels are returned as stored). */
static inline float grad_voxel_norm(const dvr_volume* v, size_t x, size_t y, size_t z, int c) {
    size_t idx = (z * (size_t)v->dims[0] * (size_t)v->dims[1] + y * (size_t)v->dims[0] + x) *
                 (size_t)v->grad_channels + (size_t)c;
    if (v->grad_dtype == DVR_U8) return (float)((const uint8_t*)v->grad_data)[idx] / 255.0f;
    if (v->grad_dtype == DVR_U16) return (float)((const uint16_t*)v->grad_data)[idx] / 65535.0f;
    return ((const float*)v->grad_data)[idx];
}

/* The gradient-volume sample of the 2D TF path, cpuraycaster.cpp:270-284:
 * channels 0..2 at the rounded voxel position (nearest) or trilinear per
 * channel (getVoxelNormalizedLinear(pos, channel), volumeram.cpp:78-97). */
static void grad_volume_sample(const dvr_volume* v, int filter, const float vp[3], float out[3]) {
    if (filter == DVR_FILTER_NEAREST) {
        size_t x = (size_t)iroundf_tgt(vp[0]), y = (size_t)iroundf_tgt(vp[1]), z = (size_t)iroundf_tgt(vp[2]);
        for (int c = 0; c < 3; ++c) out[c] = grad_voxel_norm(v, x, y, z, c);
        return;
    }
    float pa[3] = { smaxf(vp[0], 0.0f), smaxf(vp[1], 0.0f), smaxf(vp[2], 0.0f) };
    float p[3];
    size_t lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        p[i] = pa[i] - floorf(pa[i]);
        lo[i] = (size_t)pa[i];
        hi[i] = (size_t)ceilf(pa[i]);
        size_t m = (size_t)v->dims[i] - 1;
        hi[i] = hi[i] < m ? hi[i] : m;
        lo[i] = lo[i] < m ? lo[i] : m;
    }
    float ix = 1.f - p[0], iy = 1.f - p[1], iz = 1.f - p[2];
    for (int c = 0; c < 3; ++c)
        out[c] = grad_voxel_norm(v, lo[0], lo[1], lo[2], c) * ix * iy * iz
               + grad_voxel_norm(v, hi[0], lo[1], lo[2], c) * p[0] * iy * iz
               + grad_voxel_norm(v, hi[0], hi[1], lo[2], c) * p[0] * p[1] * iz
               + grad_voxel_norm(v, lo[0], hi[1], lo[2], c) * ix * p[1] * iz
               + grad_voxel_norm(v, lo[0], lo[1], hi[2], c) * ix * iy * p[2]
               + grad_voxel_norm(v, hi[0], lo[1], hi[2], c) * p[0] * iy * p[2]
               + grad_voxel_norm(v, hi[0], hi[1], hi[2], c) * p[0] * p[1] * p[2]
               + grad_voxel_norm(v, lo[0], hi[1], hi[2], c) * ix * p[1] * p[2];
}

/* Texture space -> clamped voxel position: VolumeBase::getTextureToVoxelMatrix
 * (volumebase.cpp:221-224) applied as mat4*vec4 (matrix.h:766-771: row dot
 * product starting from 0), then tgt::clamp to [0, dim-1]
 * (cpuraycaster.cpp:243). */
static inline void tex_to_voxel(const dvr_volume* v, const float s[3], float out[3]) {
    for (int i = 0; i < 3; ++i) {
        float vi = ((float)v->dims[i] * s[i]) + (-0.5f);
        out[i] = sminf(smaxf(vi, 0.0f), (float)v->dims[i] - 1.0f);
    }
}

/* Ray parameters of CPURaycaster::directRendering, cpuraycaster.cpp:219-230:
 * direction = last - first; tend = tgt::length(direction) = sqrt(lengthSq)
 * with lengthSq = ((0 + x*x) + y*y) + z*z, direction *= 1/length
 * (tgt::normalize, vector.h:802-820); a zero direction between non-zero
 * points gives tend = step/2 (one sample). */
float dvr_ray_setup(const float first[3], const float last[3], float step, float dir[3]) {
    for (int i = 0; i < 3; ++i) dir[i] = last[i] - first[i];
    int dir_zero = (dir[0] == 0.f && dir[1] == 0.f && dir[2] == 0.f);
    int last_nz = !(last[0] == 0.f && last[1] == 0.f && last[2] == 0.f);
    int first_nz = !(first[0] == 0.f && first[1] == 0.f && first[2] == 0.f);
    if (dir_zero && last_nz && first_nz) return step / 2.0f;
    float lsq = 0.0f;
    lsq += dir[0] * dir[0];
    lsq += dir[1] * dir[1];
    lsq += dir[2] * dir[2];
    float tend = sqrtf(lsq);
    float rl = 1.0f / sqrtf(lsq);
    dir[0] *= rl; dir[1] *= rl; dir[2] *= rl;
    return tend;
}

/* Sample k of a ray: t after k steps of t += step (the loop of
 * cpuraycaster.cpp:239-311), sample = first + t*dir (:240), its clamped voxel
 * position (:243) and, for the nearest filter, tgt::iround of it (:247). */
float dvr_sample_voxel(const dvr_volume* v, const float first[3], const float dir[3], float step, int32_t k,
                       float vp[3], int32_t ivp[3]) {
    float t = 0.0f;
    for (int32_t i = 0; i < k; ++i) t += step;
    float s[3] = { first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2] };
    tex_to_voxel(v, s, vp);
    for (int i = 0; i < 3; ++i) ivp[i] = iroundf_tgt(vp[i]);
    return t;
}

static inline float reconstruct(const dvr_volume* v, int filter, const float s[3]) {
    float vp[3];
    tex_to_voxel(v, s, vp);
    if (filter == DVR_FILTER_NEAREST)
        return voxel_norm(v, (size_t)iroundf_tgt(vp[0]), (size_t)iroundf_tgt(vp[1]),
                          (size_t)iroundf_tgt(vp[2]));
    return dvr_sample_linear(v, vp[0], vp[1], vp[2]);
}

static inline double reconstruct_d(const dvr_volume* v, int filter, const float s[3]) {
    float vp[3];
    tex_to_voxel(v, s, vp);
    if (filter == DVR_FILTER_NEAREST)
        return voxel_norm_d(v, (size_t)iroundf_tgt(vp[0]), (size_t)iroundf_tgt(vp[1]),
                            (size_t)iroundf_tgt(vp[2]));
    return sample_linear_d(v, vp[0], vp[1], vp[2]);
}

/* TransFuncBase::realWorldToNormalized, transfuncbase.cpp:244-259 */
static inline float rw_to_norm(float rw, float lo, float hi) {
    if (rw < lo) return 0.f;
    if (rw > hi) return 1.f;
    if (lo == hi) return 0.f;
    return (rw - lo) / (hi - lo);
}

static inline double rw_to_norm_d(double rw, double lo, double hi) {
    if (rw < lo) return 0.0;
    if (rw > hi) return 1.0;
    if (lo == hi) return 0.0;
    return (rw - lo) / (hi - lo);
}

/* ------------------------------------------------------------------------ */
/* Phong shading (GLSL semantics of SingleVolumeRaycaster), FP32 and FP64     */
/* ------------------------------------------------------------------------ */

/* calcGradientA, mod_gradients.frag:202-214: central differences at
 * +-1/dim in texture space, down-gradient, scaled by spacingRCP*0.5*rwmScale. */
static void grad_central(const dvr_volume* v, int filter, const float s[3], float g[3]) {
    float off[3], q[3], val[6];
    for (int i = 0; i < 3; ++i) off[i] = 1.f / (float)v->dims[i];
    for (int k = 0; k < 6; ++k) {
        int axis = k % 3;
        float sign = (k < 3) ? 1.f : -1.f;
        for (int i = 0; i < 3; ++i) q[i] = s[i] + ((i == axis) ? sign * off[i] : 0.0f);
        val[k] = reconstruct(v, filter, q);
    }
    for (int i = 0; i < 3; ++i) {
        float srcp = 1.f / v->spacing[i];
        g[i] = ((val[3 + i] - val[i]) * (srcp * 0.5f)) * v->rwm_scale;
    }
}

/* calcGradientAFD, mod_gradients.frag:116-126: v(s) - v(s + 1/dim) per axis,
 * scaled by spacingRCP*rwmScale. */
static void grad_forward(const dvr_volume* v, int filter, const float s[3], float g[3]) {
    float q[3];
    const float c = reconstruct(v, filter, s);
    for (int axis = 0; axis < 3; ++axis) {
        for (int i = 0; i < 3; ++i) q[i] = s[i] + ((i == axis) ? 1.f / (float)v->dims[i] : 0.0f);
        const float n = reconstruct(v, filter, q);
        g[axis] = ((c - n) * (1.f / v->spacing[axis])) * v->rwm_scale;
    }
}

/* gradient "none": (voxel.xyz - 0.5)*2 (volumeraycaster.cpp:198-199) on the
 * sample as the shader sees it: a one-channel volume texture swizzled to
 * (R, 0, 0, R) (volumegl.cpp:367) with the real-world mapping applied to all
 * four components (mod_sampler3d.frag:90-101). */
static void grad_none(const dvr_volume* v, float inten, float g[3]) {
    const float x = (inten * v->rwm_scale) + v->rwm_offset;
    const float yz = (0.0f * v->rwm_scale) + v->rwm_offset;
    g[0] = (x - 0.5f) * 2.0f;
    g[1] = (yz - 0.5f) * 2.0f;
    g[2] = (yz - 0.5f) * 2.0f;
}

/* calcGradientSobel, mod_gradients.frag:224-286.  V[n] is the shader's vNN:
 * n-1 = xg*9 + zg*3 + yg with xg = 0, 1, 2 for x = +1, 0, -1 (in units of
 * 1/dim), zg likewise for z = +1, 0, -1 and yg = 0, 1, 2 for y = -1, 0, +1.
 * The three sums are evaluated left to right as written; the factors 2 and 4
 * scale exactly. */
static void grad_sobel(const dvr_volume* v, int filter, const float s[3], float g[3]) {
    float V[28], q[3];
    const float o[3] = {1.f / (float)v->dims[0], 1.f / (float)v->dims[1], 1.f / (float)v->dims[2]};
    for (int n = 1; n <= 27; ++n) {
        if (n == 14) { V[n] = 0.f; continue; }
        const int k = n - 1, xg = k / 9, zg = (k / 3) % 3, yg = k % 3;
        const int sx = 1 - xg, sz = 1 - zg, sy = yg - 1;
        q[0] = s[0] + (sx > 0 ? o[0] : (sx < 0 ? -o[0] : 0.0f));
        q[1] = s[1] + (sy > 0 ? o[1] : (sy < 0 ? -o[1] : 0.0f));
        q[2] = s[2] + (sz > 0 ? o[2] : (sz < 0 ? -o[2] : 0.0f));
        V[n] = reconstruct(v, filter, q);
    }
    float gx = V[1] + 2.0f * V[2] + V[3] - V[19] - 2.0f * V[20] - V[21]
             + 2.0f * V[4] + 4.0f * V[5] + 2.0f * V[6] - 2.0f * V[22] - 4.0f * V[23] - 2.0f * V[24]
             + V[7] + 2.0f * V[8] + V[9] - V[25] - 2.0f * V[26] - V[27];
    float gy = -V[1] + V[3] - 2.0f * V[10] + 2.0f * V[12] - V[19] + V[21]
             - 2.0f * V[4] + 2.0f * V[6] - 4.0f * V[13] + 4.0f * V[15] - 2.0f * V[22] + 2.0f * V[24]
             - V[7] + V[9] - 2.0f * V[16] + 2.0f * V[18] - V[25] + V[27];
    float gz = V[1] + 2.0f * V[2] + V[3] + 2.0f * V[10] + 4.0f * V[11] + 2.0f * V[12] + V[19] + 2.0f * V[20] + V[21]
             - V[7] - 2.0f * V[8] - V[9] - 2.0f * V[16] - 4.0f * V[17] - 2.0f * V[18] - V[25] - 2.0f * V[26] - V[27];
    const float gs[3] = {-gx, -gy, -gz};
    for (int i = 0; i < 3; ++i) {
        float srcp = 1.f / v->spacing[i];
        g[i] = ((gs[i] * (srcp * 0.5f)) / 16.0f) * v->rwm_scale;
    }
}

/* calcGradientFiltered, mod_gradients.frag:298-314: the central gradient at
 * the sample and at its eight diagonal neighbours, blended with GLSL mix
 * (x*(1-a) + y*a). */
static void grad_filtered(const dvr_volume* v, int filter, const float s[3], float g[3]) {
    static const int diag[8][3] = {{-1, -1, -1}, {1, 1, 1}, {-1, 1, -1}, {1, -1, 1},
                                   {-1, -1, 1}, {1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
    const float d[3] = {1.f / (float)v->dims[0], 1.f / (float)v->dims[1], 1.f / (float)v->dims[2]};
    float G[9][3], q[3];
    grad_central(v, filter, s, G[0]);
    for (int k = 0; k < 8; ++k) {
        for (int i = 0; i < 3; ++i) q[i] = s[i] + (diag[k][i] > 0 ? d[i] : -d[i]);
        grad_central(v, filter, q, G[k + 1]);
    }
    for (int i = 0; i < 3; ++i) {
        const float m12 = G[1][i] * 0.5f + G[2][i] * 0.5f, m34 = G[3][i] * 0.5f + G[4][i] * 0.5f;
        const float m56 = G[5][i] * 0.5f + G[6][i] * 0.5f, m78 = G[7][i] * 0.5f + G[8][i] * 0.5f;
        const float mix0 = m12 * 0.5f + m34 * 0.5f, mix1 = m56 * 0.5f + m78 * 0.5f;
        const float m = mix0 * 0.5f + mix1 * 0.5f;
        g[i] = G[0][i] * 0.25f + m * 0.75f;
    }
}

/* CALC_GRADIENT as configured by the "gradient" property
 * (volumeraycaster.cpp:197-207). */
static void gradient_f(const dvr_volume* v, const dvr_params* P, const float s[3], float inten, float g[3]) {
    switch (P->gradient) {
    case DVR_GRAD_FORWARD: grad_forward(v, P->filter, s, g); break;
    case DVR_GRAD_NONE: grad_none(v, inten, g); break;
    case DVR_GRAD_FILTERED: grad_filtered(v, P->filter, s, g); break;
    case DVR_GRAD_SOBEL: grad_sobel(v, P->filter, s, g); break;
    default: grad_central(v, P->filter, s, g); break;
    }
}

static void gradient_d(const dvr_volume* v, int filter, const float s[3], double g[3]) {
    float off[3], q[3];
    double val[6];
    for (int i = 0; i < 3; ++i) off[i] = 1.f / (float)v->dims[i];
    for (int k = 0; k < 6; ++k) {
        int axis = k % 3;
        float sign = (k < 3) ? 1.f : -1.f;
        for (int i = 0; i < 3; ++i) q[i] = s[i] + ((i == axis) ? sign * off[i] : 0.0f);
        val[k] = reconstruct_d(v, filter, q);
    }
    for (int i = 0; i < 3; ++i)
        g[i] = (val[3 + i] - val[i]) * (0.5 / (double)v->spacing[i]) * (double)v->rwm_scale;
}

/* fixClipBorderGradient, mod_gradients.frag:39-46: normal of the proxy face
 * from the entry points two pixels to the right and two pixels up, faced
 * against the ray direction.  Out-of-image neighbours clamp to the edge. */
static void clip_border_gradient(const float* entry, int w, int h, int x, int y,
                                 const float s[3], const float dir[3], float g[3]) {
    int x2 = x + 2 < w ? x + 2 : w - 1;
    int y2 = y + 2 < h ? y + 2 : h - 1;
    const float* e0 = entry + 4 * ((size_t)y * w + x2);
    const float* e1 = entry + 4 * ((size_t)y2 * w + x);
    float v0[3], v1[3];
    for (int i = 0; i < 3; ++i) { v0[i] = e0[i] - s[i]; v1[i] = e1[i] - s[i]; }
    float l0 = sqrtf(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    float l1 = sqrtf(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    if (l0 > 0.f) { float r = 1.f / l0; for (int i = 0; i < 3; ++i) v0[i] *= r; }
    if (l1 > 0.f) { float r = 1.f / l1; for (int i = 0; i < 3; ++i) v1[i] *= r; }
    g[0] = v0[1] * v1[2] - v0[2] * v1[1];
    g[1] = v0[2] * v1[0] - v0[0] * v1[2];
    g[2] = v0[0] * v1[1] - v0[1] * v1[0];
    float dd = g[0] * dir[0] + g[1] * dir[1] + g[2] * dir[2];
    if (!(dd < 0.f)) { g[0] = -g[0]; g[1] = -g[1]; g[2] = -g[2]; }  /* faceforward */
}

/* phongShading, mod_shading.frag:134-151 (ka = kd = TF colour, ks = 1).  A
 * zero gradient gives N = 0 (GLSL leaves normalize(0) undefined). */
static void phong_f(const dvr_params* P, const float gi[3], const float pos[3],
                    const float c[3], float out[3]) {
    float gl = gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2];
    float N[3] = {0.f, 0.f, 0.f};
    if (gl > 0.f) { float r = 1.f / sqrtf(gl); for (int i = 0; i < 3; ++i) N[i] = gi[i] * r; }
    float L[3], V[3], H[3];
    for (int i = 0; i < 3; ++i) { L[i] = P->light_pos[i] - pos[i]; V[i] = P->camera_pos[i] - pos[i]; }
    float vl = sqrtf(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    float rv = 1.f / vl;
    for (int i = 0; i < 3; ++i) V[i] *= rv;
    float d = sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    for (int i = 0; i < 3; ++i) L[i] /= d;
    for (int i = 0; i < 3; ++i) H[i] = V[i] + L[i];
    float hl = sqrtf(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    float rh = 1.f / hl;
    for (int i = 0; i < 3; ++i) H[i] *= rh;
    float ndl = fmaxf(N[0] * L[0] + N[1] * L[1] + N[2] * L[2], 0.f);
    float ndh = fmaxf(N[0] * H[0] + N[1] * H[1] + N[2] * H[2], 0.f);
    float sp = powf(ndh, P->shininess);
    for (int i = 0; i < 3; ++i) {
        float amb = c[i] * P->light_ambient[i];
        float dif = (c[i] * P->light_diffuse[i]) * ndl;
        float spe = (1.0f * P->light_specular[i]) * sp;
        out[i] = ((0.0f + amb) + dif) + spe;
    }
}

/* GLSL normalize with zero -> zero (undefined in GLSL) */
static void normalize3(float v[3]) {
    float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (l2 > 0.f) { float r = 1.f / sqrtf(l2); v[0] *= r; v[1] *= r; v[2] *= r; }
    else { v[0] = v[1] = v[2] = 0.f; }
}
static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* The shading models of modules/core/glsl/modules/mod_shading.frag with the
 * arguments APPLY_SHADING passes (src/core/utils/glsl.cpp:411-434; ka = kd =
 * TF colour, ks = 1): phongShadingD :245-262, phongShadingS :188-210,
 * phongShadingDA :222-239, phongShadingDS :160-177, toonShading :273-288
 * (3 shades, unnormalized L), cookTorranceShading :303-361, orenNayarShading
 * :374-405, wardShading :419-483.  acos arguments are clamped to [-1, 1]. */
static void shade_f(const dvr_params* P, const float g[3], const float pos[3], const float c[3], float out[3]) {
    const int m = P->shading;
    if (m == DVR_SHADE_PHONG) { phong_f(P, g, pos, c, out); return; }
    float N[3] = {g[0], g[1], g[2]};
    normalize3(N);
    float L[3], V[3];
    for (int i = 0; i < 3; ++i) { L[i] = P->light_pos[i] - pos[i]; V[i] = P->camera_pos[i] - pos[i]; }
    if (m == DVR_SHADE_TOON) {
        float ndl = fmaxf(dot3(N, L), 0.f);
        for (int i = 1; i <= 3; ++i) {
            if (ndl <= (float)i / 3.0f) { ndl = (float)i / 3.0f; break; }
        }
        for (int i = 0; i < 3; ++i) out[i] = (c[i] * P->light_diffuse[i]) * ndl;
        return;
    }
    if (m <= DVR_SHADE_PHONG_DS) {
        normalize3(V);
        float d = sqrtf(dot3(L, L));
        for (int i = 0; i < 3; ++i) L[i] /= d;
        float ndl = fmaxf(dot3(N, L), 0.f);
        float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
        normalize3(H);
        float sp = powf(fmaxf(dot3(N, H), 0.f), P->shininess);
        int amb = m == DVR_SHADE_PHONG_DA;
        int dif = m == DVR_SHADE_PHONG_D || m == DVR_SHADE_PHONG_DA || m == DVR_SHADE_PHONG_DS;
        int spe = m == DVR_SHADE_PHONG_S || m == DVR_SHADE_PHONG_DS;
        for (int i = 0; i < 3; ++i) {
            float s = 0.0f;
            if (amb) s = s + c[i] * P->light_ambient[i];
            if (dif) s = s + (c[i] * P->light_diffuse[i]) * ndl;
            if (spe) s = s + (1.0f * P->light_specular[i]) * sp;
            out[i] = s;
        }
        return;
    }
    normalize3(L);
    normalize3(V);
    float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
    normalize3(H);
    if (m == DVR_SHADE_COOK_TORRANCE) {
        float rough = 1.0f - (P->shininess / 128.0f);
        float s = P->shininess / 128.0f, dd = 1.0f - s;
        float dw = sqrtf(dot3(L, L)) * 1.0f;
        float NH = dot3(N, H), NV = dot3(N, V), NL = fmaxf(0.0000001f, dot3(N, L)), VH = dot3(V, H);
        float F = 1.0f + (1.0f - 1.0f) * powf(1.0f - NV, 5.0f);
        float G1 = (2.0f * NH * NL) / VH, G2 = (2.0f * NH * NV) / VH;
        float G = fminf(1.0f, fmaxf(0.0f, fminf(G1, G2)));
        float mq = rough * rough + 1e-5f, NHq = NH * NH + 1e-5f;
        float D1 = 1.0f / (4.0f * mq * NHq * NHq);
        float D2 = expf(-(1.0f - NHq) / (mq * NHq));
        float D = D1 * D2;
        float k = ((F * D) * G) / ((3.1415926f * NV) * NL);
        for (int i = 0; i < 3; ++i) {
            float Rd = ((c[i] * P->light_diffuse[i]) * 2.0f) * NL;
            float Rs = (k * 1.0f) * P->light_specular[i];
            out[i] = (0.0f + c[i] * P->light_ambient[i]) + ((NL * dw) * (s * Rs + dd * Rd));
        }
    } else if (m == DVR_SHADE_OREN_NAYAR) {
        float rough = P->shininess / 128.0f;
        float VdotN = dot3(V, N), LdotN = dot3(L, N), NdotL = fmaxf(0.0000001f, dot3(N, L));
        float u[3], w[3];
        for (int i = 0; i < 3; ++i) { u[i] = V[i] - N[i] * VdotN; w[i] = L[i] - N[i] * LdotN; }
        float cosPhi = dot3(u, w);
        float av = acosf(fminf(fmaxf(VdotN, -1.0f), 1.0f)), al = acosf(fminf(fmaxf(LdotN, -1.0f), 1.0f));
        float alpha = fmaxf(av, al), beta = fminf(av, al);
        float mq = rough * rough;
        float A = 1.0f - 0.5f * (mq / (mq + 0.57f));
        float B = 0.45f * (mq / (mq + 0.09f));
        float r = 0.0f + (A + ((B * fmaxf(0.0f, cosPhi)) * sinf(alpha)) * tanf(beta));
        for (int i = 0; i < 3; ++i)
            out[i] = (0.0f + c[i] * P->light_ambient[i]) + ((c[i] * P->light_diffuse[i]) * NdotL) * r;
    } else {
        float ax = P->shininess / 128.0f;
        float ay = ax * 2.0f;
        float X[3] = {N[1] * 0.0f - N[2] * 0.0f, N[2] * 1.0f - N[0] * 0.0f, N[0] * 0.0f - N[1] * 1.0f};
        normalize3(X);
        float Y[3] = {N[1] * X[2] - N[2] * X[1], N[2] * X[0] - N[0] * X[2], N[0] * X[1] - N[1] * X[0]};
        normalize3(Y);
        float VdotN = dot3(V, N), LdotN = dot3(L, N), HdotN = dot3(H, N);
        float HdotX = dot3(H, X), HdotY = dot3(H, Y), NdotL = fmaxf(0.0000001f, dot3(N, L));
        ax += 0.0000001f;
        ay += 0.0000001f;
        float A = HdotX / ax; A *= A;
        float B = HdotY / ay; B *= B;
        float expo = -2.0f * ((A + B) / (1.0f + HdotN));
        float nenner = 4.0f * 3.14159f;
        nenner *= ax;
        nenner *= ay;
        nenner *= sqrtf(LdotN * VdotN);
        if (nenner < 0.0f) nenner = fminf(0.0000001f, nenner);
        else nenner = fmaxf(0.0000001f, nenner);
        float e = expf(expo) / nenner;
        for (int i = 0; i < 3; ++i) {
            float Rd = c[i] * P->light_diffuse[i];
            float Rs = (1.0f * P->light_specular[i]) * e;
            out[i] = (0.0f + c[i] * P->light_ambient[i]) + NdotL * (Rs + Rd);
        }
    }
}

void dvr_gradient(const dvr_volume* v, const dvr_params* P, const float s[3], float g[3]) {
    gradient_f(v, P, s, reconstruct(v, P->filter, s), g);
}

void dvr_shade(const dvr_params* P, const float g[3], const float pos[3], const float c[3], float out[3]) {
    shade_f(P, g, pos, c, out);
}

static void phong_d(const dvr_params* P, const double gi[3], const double pos[3],
                    const double c[3], double out[3]) {
    double gl = gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2];
    double N[3] = {0, 0, 0};
    if (gl > 0) { double r = 1.0 / sqrt(gl); for (int i = 0; i < 3; ++i) N[i] = gi[i] * r; }
    double L[3], V[3], H[3];
    for (int i = 0; i < 3; ++i) {
        L[i] = (double)P->light_pos[i] - pos[i];
        V[i] = (double)P->camera_pos[i] - pos[i];
    }
    double vl = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    for (int i = 0; i < 3; ++i) V[i] /= vl;
    double d = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    for (int i = 0; i < 3; ++i) L[i] /= d;
    for (int i = 0; i < 3; ++i) H[i] = V[i] + L[i];
    double hl = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    for (int i = 0; i < 3; ++i) H[i] /= hl;
    double ndl = fmax(N[0] * L[0] + N[1] * L[1] + N[2] * L[2], 0.0);
    double ndh = fmax(N[0] * H[0] + N[1] * H[1] + N[2] * H[2], 0.0);
    double sp = pow(ndh, (double)P->shininess);
    for (int i = 0; i < 3; ++i)
        out[i] = c[i] * P->light_ambient[i] + c[i] * P->light_diffuse[i] * ndl
               + (double)P->light_specular[i] * sp;
}

/* texToPhysical, mod_sampler3d.frag:179-181 */
static inline void tex_to_phys_f(const dvr_volume* v, const float s[3], float out[3]) {
    for (int i = 0; i < 3; ++i)
        out[i] = (((s[i] * (float)v->dims[i]) - 0.5f) * v->spacing[i]) + v->offset[i];
}

/* compositeMIDAhelper, mod_compositing.frag:112-129: maxIntensity is read,
 * never written, here. */
static void mida_helper(float res[4], float voxel, const float col[4], float max_int, float gamma) {
    float delta = 0.0f;
    if (voxel > max_int) delta = voxel - max_int;
    const float beta = 1.0f - delta * (1.0f + gamma);
    const float k = (1.0f - beta * res[3]) * col[3];
    for (int i = 0; i < 3; ++i) res[i] = beta * res[i] + k * col[i];
    res[3] = beta * res[3] + (1.0f - beta * res[3]) * col[3];
}

/* Compositing modes other than DVR (TF classification): the GLSL compositing
 * functions of mod_compositing.frag:85-213 applied where the TF alpha is > 0
 * (rc_singlevolume.frag:222-236), shading first, on the CPURaycaster sample
 * sequence; early termination for ISO/FHP/FHN (volumeraycaster.cpp:225-233,
 * mod_raysetup.frag:56-63). */
static int64_t march_modes_f(const dvr_volume* v, const dvr_tf* tf, const dvr_params* P,
                             const float first[3], const float dir[3], float tend,
                             const float* entry, int w, int h, int px, int py, float res[4]) {
    const float step = P->step;
    const int W1 = tf->width - 1;
    const int mode = P->compositing;
    const int ert = mode == DVR_COMP_ISO || mode == DVR_COMP_FHP || mode == DVR_COMP_FHN;
    float out[4] = {0.f, 0.f, 0.f, 0.f};
    float max_int = 0.0f;       /* rc_singlevolume.frag:121 */
    float t = 0.0f;
    int finished = 0;
    int64_t n = 0;
    const int cap = P->loop_cap > 0 ? P->loop_cap : 255 * 255;
    for (int loop = 0; !finished && loop < cap; ++loop) {
        ++n;
        float s[3] = { first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2] };
        const float inten = reconstruct(v, P->filter, s);
        const float rw = (inten * v->rwm_scale) + v->rwm_offset;
        const float x = rw_to_norm(rw, tf->domain_lo, tf->domain_hi);
        int idx = iroundf_tgt(x * (float)W1);
        idx = idx < 0 ? 0 : (idx > W1 ? W1 : idx);
        const uint8_t* c8 = tf->rgba8 + 4 * (size_t)idx;
        float col[4] = {(float)c8[0] / 255.f, (float)c8[1] / 255.f, (float)c8[2] / 255.f,
                        (float)c8[3] / 255.f};
        if (col[3] > 0.0f) {
            float gr[3] = {0.f, 0.f, 0.f};
            const int need_grad = P->shading != DVR_SHADE_NONE || mode == DVR_COMP_FHN;
            if (need_grad) {
                if (t == 0.0f) clip_border_gradient(entry, w, h, px, py, s, dir, gr);
                else gradient_f(v, P, s, inten, gr);
            }
            if (P->shading != DVR_SHADE_NONE) {
                float pos[3], c[3] = {col[0], col[1], col[2]}, sh[3];
                tex_to_phys_f(v, s, pos);
                shade_f(P, gr, pos, c, sh);
                col[0] = sh[0]; col[1] = sh[1]; col[2] = sh[2];
            }
            switch (mode) {
            case DVR_COMP_MIP:      /* :85-96 */
                if (rw > max_int) {
                    memcpy(out, col, sizeof out);
                    max_int = rw;
                }
                break;
            case DVR_COMP_MOP:      /* :98-107 */
                if (col[3] > out[3]) memcpy(out, col, sizeof out);
                break;
            case DVR_COMP_MIDA: {   /* :136-152 */
                col[3] = dvr_opacity_correct(col[3], step);
                if (P->gamma <= 0.0f) {
                    mida_helper(out, rw, col, max_int, P->gamma);
                } else {
                    float mip[4], mida[4];
                    memcpy(mip, out, sizeof mip);
                    memcpy(mida, out, sizeof mida);
                    if (rw > max_int) {
                        memcpy(mip, col, sizeof mip);
                        max_int = rw;
                    }
                    mida_helper(mida, rw, col, max_int, 0.0f);
                    for (int i = 0; i < 4; ++i) out[i] = P->gamma * mip[i] + (1.0f - P->gamma) * mida[i];
                }
                break;
            }
            case DVR_COMP_ISO:      /* :159-170 */
                if (col[3] >= P->iso_value - 0.02f && col[3] <= P->iso_value + 0.02f) {
                    out[0] = col[0]; out[1] = col[1]; out[2] = col[2]; out[3] = 1.0f;
                }
                break;
            case DVR_COMP_FHP:      /* :175-186 */
                if (out[0] == 0.f && out[1] == 0.f && out[2] == 0.f) {
                    out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = 1.0f;
                }
                break;
            case DVR_COMP_FHN:      /* :191-202 */
                if (out[0] == 0.f && out[1] == 0.f && out[2] == 0.f) {
                    const float l2 = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
                    const float rl = l2 > 0.f ? 1.0f / sqrtf(l2) : 0.f;
                    for (int i = 0; i < 3; ++i) out[i] = ((gr[i] * rl) / 2.0f) + 0.5f;
                    out[3] = 1.0f;
                }
                break;
            default:
                break;
            }
        }
        if (ert && out[3] >= ERT_OPACITY) {
            out[3] = 1.0f;
            finished = 1;
        }
        t += step;
        finished = finished || (t > tend);
    }
    memcpy(res, out, 4 * sizeof(float));
    return n;
}

/* ------------------------------------------------------------------------ */
/* Ray march, FP32: CPURaycaster::directRendering, cpuraycaster.cpp:210-326   */
/* ------------------------------------------------------------------------ */
static int64_t march_f(const dvr_volume* v, const dvr_tf* tf, const dvr_params* P,
                       const float first[3], const float last[3],
                       const float* entry, int w, int h, int px, int py, float res[4]) {
    const float step = P->step;
    float t = 0.0f, dir[3];
    const float tend = dvr_ray_setup(first, last, step, dir);
    const int W1 = tf->width - 1;
    float r = 0.f, g = 0.f, b = 0.f, A = 0.f;
    float last_int = 0.f;
    int finished = 0;
    int64_t n = 0;
    const int cap = P->loop_cap > 0 ? P->loop_cap : 255 * 255;
    if (P->compositing != DVR_COMP_DVR)
        return march_modes_f(v, tf, P, first, dir, tend, entry, w, h, px, py, res);
    for (int loop = 0; !finished && loop < cap; ++loop) {
        ++n;
        float s[3] = { first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2] };
        float inten = reconstruct(v, P->filter, s);
        float rw = (inten * v->rwm_scale) + v->rwm_offset;
        float cr, cg, cb, ca;
        if (P->classification == DVR_CLASS_PREINT) {
            float x = rw_to_norm(rw, tf->domain_lo, tf->domain_hi);
            int N = P->preint_n;
            int i1 = iroundf_tgt(last_int * (float)(N - 1));
            int i2 = iroundf_tgt(x * (float)(N - 1));
            i1 = i1 < 0 ? 0 : (i1 > N - 1 ? N - 1 : i1);
            i2 = i2 < 0 ? 0 : (i2 > N - 1 ? N - 1 : i2);
            const float* c = P->preint_table + 4 * ((size_t)i2 * N + i1);
            cr = c[0]; cg = c[1]; cb = c[2]; ca = c[3];
            last_int = x;
        } else if (P->classification == DVR_CLASS_TF2D) {
            /* cpuraycaster.cpp:270-290 with apply2DTF :334-338: intensity and
             * gradient magnitude (tgt::length) each mapped by its domain, the
             * texel at the truncated positions, then opacity correction */
            float vp[3], gv[3];
            tex_to_voxel(v, s, vp);
            grad_volume_sample(v, P->filter, vp, gv);
            float lsq = 0.0f;
            lsq += gv[0] * gv[0];
            lsq += gv[1] * gv[1];
            lsq += gv[2] * gv[2];
            const float xi = rw_to_norm(rw, tf->domain2d[0], tf->domain2d[1]);
            const float xg = rw_to_norm(sqrtf(lsq), tf->domain2d[2], tf->domain2d[3]);
            const size_t ix = (size_t)(xi * (float)(tf->width2d - 1));
            const size_t iy = (size_t)(xg * (float)(tf->height2d - 1));
            const float* c = tf->rgba2d + 4 * (iy * (size_t)tf->width2d + ix);
            cr = c[0]; cg = c[1]; cb = c[2];
            ca = dvr_opacity_correct(c[3], step);
        } else {
            float x = rw_to_norm(rw, tf->domain_lo, tf->domain_hi);
            int idx = iroundf_tgt(x * (float)W1);
            idx = idx < 0 ? 0 : (idx > W1 ? W1 : idx);
            const uint8_t* c8 = tf->rgba8 + 4 * (size_t)idx;
            cr = (float)c8[0] / 255.f;
            cg = (float)c8[1] / 255.f;
            cb = (float)c8[2] / 255.f;
            ca = (float)c8[3] / 255.f;
            ca = dvr_opacity_correct(ca, step);
        }
        if (ca > 0.0f) {
            if (P->shading != DVR_SHADE_NONE) {
                float gr[3], pos[3], c[3] = {cr, cg, cb}, sh[3];
                if (t == 0.0f) clip_border_gradient(entry, w, h, px, py, s, dir, gr);
                else gradient_f(v, P, s, inten, gr);
                tex_to_phys_f(v, s, pos);
                shade_f(P, gr, pos, c, sh);
                cr = sh[0]; cg = sh[1]; cb = sh[2];
            }
            float wgt = (1.0f - A) * ca;
            r = r + wgt * cr;
            g = g + wgt * cg;
            b = b + wgt * cb;
            A = A + (1.0f - A) * ca;
        }
        if (A >= ERT_OPACITY) {
            A = 1.0f;
            finished = 1;
        }
        t += step;
        finished = finished || (t > tend);
    }
    res[0] = r; res[1] = g; res[2] = b; res[3] = A;
    return n;
}

/* Same sample positions (float recurrence), everything else in double. */
static int64_t march_d(const dvr_volume* v, const dvr_tf* tf, const dvr_params* P,
                       const float first[3], const float last[3],
                       const float* entry, int w, int h, int px, int py, double res[4]) {
    const float step = P->step;
    float t = 0.0f, dir[3];
    const float tend = dvr_ray_setup(first, last, step, dir);
    if (P->compositing != DVR_COMP_DVR) {
        /* the FP64 envelope is defined for DVR only; other modes are the FP32 result */
        float rf[4];
        const int64_t k = march_modes_f(v, tf, P, first, dir, tend, entry, w, h, px, py, rf);
        for (int i = 0; i < 4; ++i) res[i] = rf[i];
        return k;
    }
    if ((P->shading != DVR_SHADE_NONE && P->shading != DVR_SHADE_PHONG) ||
        (P->shading != DVR_SHADE_NONE && P->gradient != DVR_GRAD_CENTRAL) ||
        P->classification == DVR_CLASS_TF2D) {
        /* and for Phong on central differences with 1D classification; the
         * other shading models, gradients and the 2D TF are the FP32 result */
        float rf[4];
        const int64_t k = march_f(v, tf, P, first, last, entry, w, h, px, py, rf);
        for (int i = 0; i < 4; ++i) res[i] = rf[i];
        return k;
    }
    const int W1 = tf->width - 1;
    const double e = (double)step * 200.0;
    double r = 0, g = 0, b = 0, A = 0, last_int = 0;
    int finished = 0;
    int64_t n = 0;
    const int cap = P->loop_cap > 0 ? P->loop_cap : 255 * 255;
    for (int loop = 0; !finished && loop < cap; ++loop) {
        ++n;
        float s[3] = { first[0] + t * dir[0], first[1] + t * dir[1], first[2] + t * dir[2] };
        double inten = reconstruct_d(v, P->filter, s);
        double rw = inten * (double)v->rwm_scale + (double)v->rwm_offset;
        double x = rw_to_norm_d(rw, tf->domain_lo, tf->domain_hi);
        double cr, cg, cb, ca;
        if (P->classification == DVR_CLASS_PREINT) {
            int N = P->preint_n;
            int i1 = iround_d(last_int * (N - 1));
            int i2 = iround_d(x * (N - 1));
            i1 = i1 < 0 ? 0 : (i1 > N - 1 ? N - 1 : i1);
            i2 = i2 < 0 ? 0 : (i2 > N - 1 ? N - 1 : i2);
            const float* c = P->preint_table + 4 * ((size_t)i2 * N + i1);
            cr = c[0]; cg = c[1]; cb = c[2]; ca = c[3];
            last_int = x;
        } else {
            int idx = iround_d(x * W1);
            idx = idx < 0 ? 0 : (idx > W1 ? W1 : idx);
            const uint8_t* c8 = tf->rgba8 + 4 * (size_t)idx;
            cr = c8[0] / 255.0; cg = c8[1] / 255.0; cb = c8[2] / 255.0; ca = c8[3] / 255.0;
            ca = 1.0 - pow(1.0 - ca, e);
        }
        if (ca > 0.0) {
            if (P->shading == DVR_SHADE_PHONG) {
                double gr[3], pos[3], c[3] = {cr, cg, cb}, sh[3];
                if (t == 0.0f) {
                    float gf[3];
                    clip_border_gradient(entry, w, h, px, py, s, dir, gf);
                    for (int i = 0; i < 3; ++i) gr[i] = gf[i];
                } else {
                    gradient_d(v, P->filter, s, gr);
                }
                for (int i = 0; i < 3; ++i)
                    pos[i] = (((double)s[i] * v->dims[i]) - 0.5) * v->spacing[i] + v->offset[i];
                phong_d(P, gr, pos, c, sh);
                cr = sh[0]; cg = sh[1]; cb = sh[2];
            }
            double wgt = (1.0 - A) * ca;
            r += wgt * cr; g += wgt * cg; b += wgt * cb;
            A += (1.0 - A) * ca;
        }
        if (A >= 0.95) { A = 1.0; finished = 1; }
        t += step;
        finished = finished || (t > tend);
    }
    res[0] = r; res[1] = g; res[2] = b; res[3] = A;
    return n;
}

/* CPURaycaster::process pixel loop, cpuraycaster.cpp:173-193: rows serial,
 * pixels of a row in an OpenMP parallel for with dynamic schedule; a pixel is
 * background only if entry and exit are both the zero vec4 (:184). */
int64_t dvr_cpu_render(const dvr_volume* vol, const dvr_tf* tf, const dvr_params* prm,
                       const float* entry, const float* exit_, int32_t w, int32_t h,
                       float* out, int32_t nthreads) {
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    for (int y = 0; y < h; ++y) {
        int64_t row = 0;
        #pragma omp parallel for schedule(dynamic) reduction(+:row)
        for (int x = 0; x < w; ++x) {
            size_t p = (size_t)y * w + x;
            const float* f = entry + 4 * p;
            const float* l = exit_ + 4 * p;
            float* o = out + 4 * p;
            int bg = (f[0] == 0.f && f[1] == 0.f && f[2] == 0.f && f[3] == 0.f &&
                      l[0] == 0.f && l[1] == 0.f && l[2] == 0.f && l[3] == 0.f);
            if (bg) {
                o[0] = o[1] = o[2] = o[3] = 0.f;
            } else {
                row += march_f(vol, tf, prm, f, l, entry, w, h, x, y, o);
            }
        }
        total += row;
    }
    return total;
}

int64_t dvr_fp64_render(const dvr_volume* vol, const dvr_tf* tf, const dvr_params* prm,
                        const float* entry, const float* exit_, int32_t w, int32_t h,
                        double* out, int32_t nthreads) {
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 16) reduction(+:total)
    for (long long p = 0; p < (long long)w * h; ++p) {
        int x = (int)(p % w), y = (int)(p / w);
        const float* f = entry + 4 * p;
        const float* l = exit_ + 4 * p;
        double* o = out + 4 * p;
        int bg = (f[0] == 0.f && f[1] == 0.f && f[2] == 0.f && f[3] == 0.f &&
                  l[0] == 0.f && l[1] == 0.f && l[2] == 0.f && l[3] == 0.f);
        if (bg) {
            o[0] = o[1] = o[2] = o[3] = 0.0;
        } else {
            total += march_d(vol, tf, prm, f, l, entry, w, h, x, y, o);
        }
    }
    return total;
}
