/*
 * dvr_oracle.h -- CPU oracle for the direct-volume-rendering ray march.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (voreen_amd, libhipdvr.so) never links, loads or calls it.
 *
 * It restates, in plain C, the semantics of Voreen's CPURaycaster and the
 * pieces it depends on (reference paths relative to voreen/):
 *
 *   dvr_cpu_render        modules/base/processors/render/cpuraycaster.cpp:90-332
 *                         (FP32, bit-faithful op order; OpenMP schedule(dynamic)
 *                         per image row like :173-193)
 *   dvr_fp64_render       same sample positions, everything after the sample
 *                         position in double precision (tolerance envelope)
 *   dvr_tf_build          src/core/datastructures/transfunc/1d/transfunc1d.cpp:85-116
 *                         + 1dkeys/transfunc1dkeys.cpp:119-154
 *                         + transfuncbase.cpp:177-210 (alpha modes)
 *   dvr_preint_build      src/core/datastructures/transfunc/1d/preintegrationtable.cpp:67-204
 *   dvr_entry_exit        analytic cube-proxy entry/exit points (the analogue of
 *                         modules/base/processors/entryexitpoints/meshentryexitpoints.cpp:188-273
 *                         for a bounding-box proxy, with the camera-inside fill of
 *                         modules/base/glsl/eep_insidevolume.frag:44-76)
 *   Phong (shading=1)     modules/core/glsl/modules/mod_gradients.frag:39-46,202-214,
 *                         mod_shading.frag:134-151, mod_sampler3d.frag:179-181,
 *                         modules/base/glsl/rc_singlevolume.frag:222-235
 *   compositing modes     MIP/MOP/MIDA/ISO/FHP/FHN of
 *                         modules/core/glsl/modules/mod_compositing.frag:85-213 as
 *                         invoked by rc_singlevolume.frag:200-266 (composite only
 *                         where the TF alpha is > 0; early termination only for
 *                         DVR/ISO/FHP/FHN, src/core/processors/volumeraycaster.cpp:
 *                         225-233), on the CPURaycaster sample sequence (FP32 only)
 *
 * Parity pinning of every function is listed in the header of dvr_oracle.c.
 */
#ifndef DVR_ORACLE_H
#define DVR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DVR_U8 = 0, DVR_U16 = 1, DVR_F32 = 2 };  /* DVR_F32: gradient volumes only */
enum { DVR_FILTER_NEAREST = 0, DVR_FILTER_LINEAR = 1 };
enum { DVR_CLASS_TF = 0, DVR_CLASS_PREINT = 1, DVR_CLASS_TF2D = 2 };
/* VolumeRaycaster "gradient" option keys (volumeraycaster.cpp:126-131,197-207) */
enum { DVR_GRAD_CENTRAL = 0, DVR_GRAD_FORWARD = 1, DVR_GRAD_NONE = 2, DVR_GRAD_FILTERED = 3,
       DVR_GRAD_SOBEL = 4 };
/* shading models, glsl.cpp:397-434 option keys in this order */
enum { DVR_SHADE_NONE = 0, DVR_SHADE_PHONG = 1, DVR_SHADE_PHONG_D = 2, DVR_SHADE_PHONG_S = 3,
       DVR_SHADE_PHONG_DA = 4, DVR_SHADE_PHONG_DS = 5, DVR_SHADE_TOON = 6, DVR_SHADE_COOK_TORRANCE = 7,
       DVR_SHADE_OREN_NAYAR = 8, DVR_SHADE_WARD = 9 };
/* compositing modes, mod_compositing.frag:47-230 / volumeraycaster.cpp:250-271 */
enum { DVR_COMP_DVR = 0, DVR_COMP_MIP = 1, DVR_COMP_MOP = 2, DVR_COMP_MIDA = 3, DVR_COMP_ISO = 4,
       DVR_COMP_FHP = 5, DVR_COMP_FHN = 6 };
/* TransFuncBase::AlphaMode (include/voreen/core/datastructures/transfunc/transfuncbase.h:165-169) */
enum { DVR_ALPHA_ZERO = 0, DVR_ALPHA_USE = 1, DVR_ALPHA_ONE = 2 };

typedef struct {
    int32_t dims[3];
    int32_t dtype;          /* DVR_U8 / DVR_U16 */
    const void* data;       /* x-fastest linear: z*X*Y + y*X + x */
    float rwm_scale;        /* RealWorldMapping: rw = n*scale + offset */
    float rwm_offset;
    float spacing[3];
    float offset[3];
    /* gradient volume of the 2D TF path (cpuraycaster.cpp:134-141,270-288): same
     * dims, grad_channels (3 or 4) interleaved channels of grad_dtype per voxel */
    const void* grad_data;
    int32_t grad_dtype;     /* DVR_U8 / DVR_U16 / DVR_F32 */
    int32_t grad_channels;
} dvr_volume;

typedef struct {
    const uint8_t* rgba8;   /* width x RGBA8 */
    int32_t width;
    float domain_lo;
    float domain_hi;
    /* 2D TF texture (TransFunc2D::getTexture, float RGBA, texel (x, y) at
     * y*width2d + x) and its per-dimension domains {intensity lo, hi, gradient
     * magnitude lo, hi} (transfunc2d.cpp:229-231) */
    const float* rgba2d;
    int32_t width2d;
    int32_t height2d;
    float domain2d[4];
} dvr_tf;

typedef struct {
    float step;             /* texture-space sampling step */
    int32_t filter;
    int32_t classification;
    const float* preint_table;  /* N*N*RGBA float, index [fe*N + fs] */
    int32_t preint_n;
    int32_t shading;
    float light_pos[3];     /* physical space */
    float light_ambient[3];
    float light_diffuse[3];
    float light_specular[3];
    float shininess;
    float camera_pos[3];    /* physical space */
    int32_t loop_cap;       /* 255*255 in the reference */
    int32_t compositing;    /* DVR_COMP_*; modes other than DVR need TF classification */
    float gamma;            /* MIDA gammaValue, [-1, 1] (volumeraycaster.cpp:67) */
    float iso_value;        /* ISO isoValue (volumeraycaster.cpp:57) */
    int32_t gradient;       /* DVR_GRAD_*: gradient of shading and FHN */
} dvr_params;

typedef struct {
    float intensity;
    int32_t split;
    uint8_t color_l[4];
    uint8_t color_r[4];
} dvr_tf_key;

typedef struct {
    float eye[3];
    float focus[3];
    float up[3];
    float fovy_deg;
    float near_dist;
} dvr_camera;

/* ---- seeded PRNG (SplitMix64) ---- */
uint64_t dvr_splitmix64(uint64_t* state);
double dvr_uniform01(uint64_t* state);

/* ---- transfer function ---- */
void dvr_tf_build(const dvr_tf_key* keys, int32_t nkeys, int32_t width, float gamma,
                  float thr_lo, float thr_hi, int32_t alpha_mode, uint8_t* out_rgba8);
/* opacity correction of one TF-mode sample, cpuraycaster.cpp:291 */
float dvr_opacity_correct(float alpha, float step);

/* ---- pre-integration table, preintegrationtable.cpp:67-204 ---- */
void dvr_preint_build(const dvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo,
                      float thr_hi, int32_t alpha_mode, int32_t n, float step,
                      int32_t use_integral, float* out_rgba);

/* ---- entry/exit points ---- */
void dvr_entry_exit(const dvr_camera* cam, const float llf[3], const float urb[3],
                    int32_t w, int32_t h, int32_t quantize16, float* entry, float* exit_);

/* ---- renderers; both return the number of samples (loop iterations) ---- */
int64_t dvr_cpu_render(const dvr_volume* vol, const dvr_tf* tf, const dvr_params* prm,
                       const float* entry, const float* exit_, int32_t w, int32_t h,
                       float* out_rgba, int32_t nthreads);
int64_t dvr_fp64_render(const dvr_volume* vol, const dvr_tf* tf, const dvr_params* prm,
                        const float* entry, const float* exit_, int32_t w, int32_t h,
                        double* out_rgba, int32_t nthreads);

/* one shading evaluation (params' model, light, camera, shininess) of TF
 * colour c with gradient g at physical position pos (for unit tests) */
void dvr_shade(const dvr_params* prm, const float g[3], const float pos[3], const float c[3], float out[3]);
/* Ray setup of directRendering: normalized direction, returns tend. */
float dvr_ray_setup(const float first[3], const float last[3], float step, float dir[3]);
/* Sample k of a ray: returns t; vp = clamped voxel position, ivp = its tgt::iround. */
float dvr_sample_voxel(const dvr_volume* vol, const float first[3], const float dir[3], float step, int32_t k,
                       float vp[3], int32_t ivp[3]);
/* CALC_GRADIENT of prm->gradient at texture-space position s (FP32). */
void dvr_gradient(const dvr_volume* vol, const dvr_params* prm, const float s[3], float g[3]);

/* single trilinear reconstruction, volumeram.cpp:78-97 (for unit tests) */
float dvr_sample_linear(const dvr_volume* vol, float vx, float vy, float vz);

/* ---- stepping network oracle (FP64), TEST INFRASTRUCTURE ONLY ----
 * 16 -> 64 (GELU) -> 64 (GELU) -> 2; params as the library's flat FP32
 * layout W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2].  With
 * quantize != 0 the weight matrices are first rounded to BF16 (nearest even),
 * as the GPU kernels use them; activations stay FP64.  Loss per row:
 * BCE-with-logits(y0, t0) + t2*(y1 - t1)^2, mean over rows. */
#define DVR_MLP_PARAMS (16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2)
void dvr_mlp_forward(const float* params, int32_t quantize, const float* x, int32_t rows, double* y);
double dvr_mlp_loss_grad(const double* params, int32_t quantize, const float* x, const float* t,
                         int32_t rows, double* grad);

#ifdef __cplusplus
}
#endif
#endif
