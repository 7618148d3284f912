/*
 * hipdvr.h -- C-ABI of the MI355X (gfx950) direct-volume-rendering ray march.
 *
 * This is the boundary between a Voreen processor (HIPRaycaster, the drop-in
 * for the ray-march path of CPURaycaster) and the HIP kernels.  Plain C types
 * only: pointers, sizes, PODs.  No C++ exception crosses it; every entry point
 * returns a status (HIPDVR_OK == 0) and the message of the last failure is
 * kept per context (hipdvr_last_error).
 *
 * Reference interfaces each entry point replaces (paths relative to voreen/):
 *
 *   hipdvr_create / hipdvr_destroy
 *       Processor::initialize / deinitialize of the raycaster
 *       (include/voreen/core/processors/processor.h:401-445,
 *        modules/base/processors/render/cpuraycaster.cpp:86-88).
 *   hipdvr_volume_create / hipdvr_volume_destroy
 *       VolumeRAM data handed to the raycaster: getRepresentation<VolumeRAM>(),
 *       getDimensions(), getBaseType(), getRealWorldMapping(), spacing/offset
 *       (cpuraycaster.cpp:99-101,215; src/core/datastructures/volume/volumebase.cpp:
 *       158-172,221-229,375-385).  Uploaded once per INVALID_PROGRAM
 *       (src/core/processors/volumeraycaster.cpp:44).
 *   hipdvr_tf_create / hipdvr_tf_destroy
 *       the RGBA8 1D lookup texture and domain of tFunc1DKeys_
 *       (cpuraycaster.cpp:130,154,266-268,328-332;
 *        src/core/datastructures/transfunc/1d/transfunc1d.cpp:85-116,217-219).
 *   hipdvr_tf_from_keys
 *       TransFunc1D::createTextureData + TransFunc1DKeys::getMappingForValueUByte
 *       + TransFuncBase::applyAlpha (transfunc1d.cpp:85-116,
 *        1d/1dkeys/transfunc1dkeys.cpp:119-154, transfuncbase.cpp:177-210).
 *   hipdvr_tf_set_preint
 *       PreIntegrationTable::classify lookup table for the "pre-integrated-cpu" /
 *       "pre-integrated-approximation" classification modes
 *       (cpuraycaster.cpp:114-122,257-263;
 *        src/core/datastructures/transfunc/1d/preintegrationtable.cpp:256-279).
 *   hipdvr_preint_build
 *       PreIntegrationTable::computeTable (preintegrationtable.cpp:67-204).
 *   hipdvr_tf_preint_device
 *       the same table computed on the device and attached to the TF: the
 *       "pre-integrated-gpu" classification mode (classificationmodes.cpp:84-90,
 *       PreIntegrationTable::computeTable's GPU path preintegrationtable.cpp:
 *       206-254 in the reference; here the exact/integral CPU algorithm runs
 *       one lane per table entry).
 *   hipdvr_tf2d_create
 *       the float RGBA texture and domains of tFunc2DPrimitives_ for the
 *       TFT_2DPRIMITIVES path (cpuraycaster.cpp:133-148,334-338;
 *       src/core/datastructures/transfunc/2d/transfunc2d.cpp:229-231).
 *   hipdvr_volume_set_gradient
 *       the RGB(A) VolumeRAM of gradientVolumePort_ that the 2D TF path samples
 *       (cpuraycaster.cpp:46,134-141,270-284).
 *   hipdvr_render
 *       the pixel loop and ray march of CPURaycaster::process/directRendering
 *       (cpuraycaster.cpp:157-208 with :210-326), host entry/exit/output buffers
 *       as downloaded/uploaded at :163-169,197-199.
 *   hipdvr_render_device
 *       the same on device-resident buffers, asynchronous on the context stream.
 *   hipdvr_entry_exit_device
 *       MeshEntryExitPoints::process for the bounding-box proxy, including the
 *       camera-inside near-plane fill
 *       (modules/base/processors/entryexitpoints/meshentryexitpoints.cpp:188-273,
 *        modules/base/glsl/eep_insidevolume.frag:44-76).
 *   hipdvr_buffer_* / hipdvr_memcpy_*
 *       device staging used by callers without their own allocator (the
 *       reference keeps these as tgt::Texture objects of the render ports,
 *       include/voreen/core/ports/renderport.h:116).
 */
#ifndef HIPDVR_H
#define HIPDVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPDVR_ABI_VERSION 2

/* status codes */
enum {
    HIPDVR_OK = 0,
    HIPDVR_ERR_INVALID = 1,     /* bad argument or handle */
    HIPDVR_ERR_HIP = 2,         /* HIP runtime failure (no device, launch error, ...) */
    HIPDVR_ERR_NOMEM = 3,       /* device or host allocation failed */
    HIPDVR_ERR_UNSUPPORTED = 4  /* option combination not implemented */
};

/* voxel types: VolumeElement<uint8_t>/<uint16_t>; float channels for gradient
 * volumes only (getTypeAsFloat(float), volumeelement.h:1243-1245) */
enum { HIPDVR_U8 = 0, HIPDVR_U16 = 1, HIPDVR_F32 = 2 };

/* textureFilterMode_ option keys "nearest"/"linear" (cpuraycaster.cpp:62-64) */
enum { HIPDVR_FILTER_NEAREST = 0, HIPDVR_FILTER_LINEAR = 1 };

/* classification option keys (src/core/utils/classificationmodes.cpp:84-90) */
enum {
    HIPDVR_CLASS_TF = 0,        /* "transfer-function" */
    HIPDVR_CLASS_PREINT = 1,    /* "pre-integrated-cpu" / "-approximation" (table given) */
    HIPDVR_CLASS_TF2D = 2       /* "transfer-function" with tFuncType tFunc2DPrimitives: a 2D
                                   TF over intensity and gradient-volume magnitude
                                   (cpuraycaster.cpp:133-148,270-290) */
};

/* gradient option keys (volumeraycaster.cpp:126-131, CALC_GRADIENT :197-207;
 * mod_gradients.frag): the gradient used by shading and FHN compositing */
enum {
    HIPDVR_GRADIENT_CENTRAL = 0,   /* "central-differences" (the default) */
    HIPDVR_GRADIENT_FORWARD = 1,   /* "forward-differences" */
    HIPDVR_GRADIENT_NONE = 2,      /* "none": (voxel.xyz - 0.5) * 2 */
    HIPDVR_GRADIENT_FILTERED = 3,  /* "filtered" */
    HIPDVR_GRADIENT_SOBEL = 4      /* "sobel" */
};

/* compositing option keys (volumeraycaster.cpp:379-388); the modes other
 * than DVR follow modules/core/glsl/modules/mod_compositing.frag:85-213 and
 * need transfer-function classification */
enum {
    HIPDVR_COMPOSITE_DVR = 0,   /* "dvr" */
    HIPDVR_COMPOSITE_MIP = 1,   /* "mip" */
    HIPDVR_COMPOSITE_MOP = 2,   /* "mop" */
    HIPDVR_COMPOSITE_MIDA = 3,  /* "mida" (gamma) */
    HIPDVR_COMPOSITE_ISO = 4,   /* "iso" (iso_value) */
    HIPDVR_COMPOSITE_FHP = 5,   /* "fhp" first-hit position */
    HIPDVR_COMPOSITE_FHN = 6    /* "fhn" first-hit normal */
};

/* shading option keys (src/core/utils/glsl.cpp:397-409) */
enum {
    HIPDVR_SHADE_NONE = 0,                   /* "none" */
    HIPDVR_SHADE_PHONG = 1,                  /* "phong" (mod_shading.frag:134-151) */
    HIPDVR_SHADE_PHONG_DIFFUSE = 2,          /* "phong-diffuse" */
    HIPDVR_SHADE_PHONG_SPECULAR = 3,         /* "phong-specular" */
    HIPDVR_SHADE_PHONG_DIFFUSE_AMBIENT = 4,  /* "phong-diffuse-ambient" */
    HIPDVR_SHADE_PHONG_DIFFUSE_SPECULAR = 5, /* "phong-diffuse-specular" */
    HIPDVR_SHADE_TOON = 6,                   /* "toon" (3 shades) */
    HIPDVR_SHADE_COOK_TORRANCE = 7,          /* "cook-torrance" */
    HIPDVR_SHADE_OREN_NAYAR = 8,             /* "oren-nayar" */
    HIPDVR_SHADE_WARD = 9                    /* "ward" */
};

/* TransFuncBase::AlphaMode (transfuncbase.h:165-169) */
enum { HIPDVR_ALPHA_ZERO = 0, HIPDVR_ALPHA_USE = 1, HIPDVR_ALPHA_ONE = 2 };

/* kernel selection */
enum {
    HIPDVR_KERNEL_AUTO = 0,      /* fastest exact kernel for the options */
    HIPDVR_KERNEL_REFERENCE = 1, /* linear layout, one lane per ray, no skipping */
    HIPDVR_KERNEL_BRICKED = 2,   /* bricked layout + exact empty-space skipping */
    HIPDVR_KERNEL_CELL = 3,      /* corner-replicated cell records + exact empty-space
                                    skipping; linear filter, TF classification */
    HIPDVR_KERNEL_SLAB = 4       /* the linear volume staged into LDS per image tile, one
                                    slab of planes at a time + exact empty-space skipping;
                                    linear filter, TF classification */
};

typedef struct hipdvr_ctx hipdvr_ctx;
typedef struct hipdvr_volume hipdvr_volume;
typedef struct hipdvr_tf hipdvr_tf;

typedef struct hipdvr_volume_desc {
    uint32_t dims[3];       /* x, y, z; voxels are x-fastest: z*X*Y + y*X + x */
    int32_t dtype;          /* HIPDVR_U8 / HIPDVR_U16 */
    float rwm_scale;        /* RealWorldMapping: rw = normalized*scale + offset */
    float rwm_offset;
    float spacing[3];       /* physical voxel size (Phong only) */
    float offset[3];        /* physical position of voxel 0 (Phong only) */
    uint32_t brick_log2;    /* 3, 4 or 5 (8^3, 16^3, 32^3 bricks); 0 = default 4 */
} hipdvr_volume_desc;

typedef struct hipdvr_params {
    float step;             /* texture-space step, 1/(max(dim)*samplingRate)
                               (cpuraycaster.cpp:104) */
    int32_t filter;         /* HIPDVR_FILTER_* */
    int32_t classification; /* HIPDVR_CLASS_* */
    int32_t compositing;    /* HIPDVR_COMPOSITE_DVR */
    int32_t shading;        /* HIPDVR_SHADE_* */
    float light_pos[3];     /* physical space (VolumeRenderer defaults
                               2.3,1.5,1.5; src/core/processors/volumerenderer.cpp:43) */
    float light_ambient[3];
    float light_diffuse[3];
    float light_specular[3];
    float shininess;
    float camera_pos[3];    /* physical space */
    int32_t loop_cap;       /* 0 = 255*255 (cpuraycaster.cpp:239) */
    int32_t kernel;         /* HIPDVR_KERNEL_* */
    int32_t batch;          /* cell kernel: samples fetched together per ray (1, 2 or 4,
                               8 runs as 4; 0 = 2 shaded, 4 unshaded) */
    float gamma;            /* MIDA gammaValue in [-1, 1] (volumeraycaster.cpp:67) */
    float iso_value;        /* ISO isoValue in [0, 1] (volumeraycaster.cpp:57) */
    int32_t exact_shading;  /* Phong in the cell kernel: 1 = bit-identical to the reference
                               kernel; 0 = interior gradient from the sample's own records and
                               hardware rsq/exp2/log2 lighting (within the oracle tolerance) */
    int32_t gradient;       /* HIPDVR_GRADIENT_*; other than central differences the shaded
                               (or FHN) render runs on the reference kernel */
} hipdvr_params;

typedef struct hipdvr_stats {
    uint64_t rays;          /* non-background pixels */
    uint64_t samples;       /* ray-march loop iterations (the reference's steps) */
    uint64_t fetched;       /* samples whose volume value was reconstructed */
    uint64_t shaded;        /* samples shaded (Phong) */
    uint64_t lane_slots;    /* 64 x loop iterations of each wavefront (SIMD-efficiency
                               diagnostic: samples / lane_slots) */
    uint64_t shade_slots;   /* 64 x iterations in which any lane of the wavefront shaded */
    float kernel_ms;        /* HIP-event time of the march kernel on the ctx stream */
} hipdvr_stats;

typedef struct hipdvr_tf_key {
    float intensity;        /* TransFuncMappingKey (1d/1dkeys/utils/transfuncmappingkey.cpp:30-106) */
    int32_t split;
    uint8_t color_l[4];
    uint8_t color_r[4];
} hipdvr_tf_key;

typedef struct hipdvr_camera {
    float eye[3];           /* world == physical space (identity transformation) */
    float focus[3];
    float up[3];
    float fovy_deg;         /* tgt::Camera (ext/tgt/camera.cpp:203-214): clamped to [6, 175], spans the
                             * image height when w >= h, else its width (frustum ratio 1) */
    float near_dist;
} hipdvr_camera;

/* ---- context ---- */
int hipdvr_create(hipdvr_ctx** out, int device);
int hipdvr_destroy(hipdvr_ctx* ctx);
const char* hipdvr_last_error(const hipdvr_ctx* ctx);
int hipdvr_abi_version(void);
/* SHA-256 (hex) of the sources and public headers the library was built from
 * (no counterpart in the reference): lets a caller that loads a prebuilt
 * library check that it matches the sources shipped with it. */
const char* hipdvr_source_digest(void);
/* use a caller-owned HIP stream (hipStream_t as void*; NULL = context stream) */
int hipdvr_set_stream(hipdvr_ctx* ctx, void* stream);
int hipdvr_synchronize(hipdvr_ctx* ctx);

/* ---- volume ---- */
int hipdvr_volume_create(hipdvr_ctx* ctx, const hipdvr_volume_desc* desc,
                         const void* host_voxels, hipdvr_volume** out);
int hipdvr_volume_create_device(hipdvr_ctx* ctx, const hipdvr_volume_desc* desc,
                                const void* device_voxels, hipdvr_volume** out);
int hipdvr_volume_destroy(hipdvr_volume* vol);
/* device pointer of the linear (x-fastest) copy */
int hipdvr_volume_device_ptr(const hipdvr_volume* vol, const void** out);
/* Attach the gradient volume of the 2D TF path: the volume's dims, `channels`
 * (3 or 4) interleaved channels of dtype (HIPDVR_U8 / U16 / F32) per voxel,
 * x-fastest; channels 0..2 are read (cpuraycaster.cpp:46 gradientVolumePort_,
 * checks :134-141, sampling :270-284). */
int hipdvr_volume_set_gradient(hipdvr_volume* vol, const void* host_data, int32_t dtype, uint32_t channels);

/* ---- transfer function ---- */
int hipdvr_tf_from_keys(const hipdvr_tf_key* keys, int32_t nkeys, int32_t width, float gamma,
                        float thr_lo, float thr_hi, int32_t alpha_mode, uint8_t* out_rgba8);
int hipdvr_tf_create(hipdvr_ctx* ctx, const uint8_t* rgba8, uint32_t width, float domain_lo,
                     float domain_hi, hipdvr_tf** out);
int hipdvr_tf_destroy(hipdvr_tf* tf);
/* A 2D transfer function: the float RGBA texture of TransFunc2DPrimitives
 * (texel (x, y) at rgba[4*(y*width + x)], x = intensity, y = gradient
 * magnitude) with domain = {intensity lo, hi, magnitude lo, hi}
 * (transfunc2d.cpp:229-231).  Rendered with HIPDVR_CLASS_TF2D; the opacity
 * correction is applied per step on the host, as cpuraycaster.cpp:289 does
 * per sample. */
int hipdvr_tf2d_create(hipdvr_ctx* ctx, const float* rgba, uint32_t width, uint32_t height,
                       const float domain[4], hipdvr_tf** out);
/* N x N RGBA float table, element [fe_index*N + fs_index] */
int hipdvr_tf_set_preint(hipdvr_tf* tf, const float* table_rgba, uint32_t n);
int hipdvr_preint_build(const hipdvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo,
                        float thr_hi, int32_t alpha_mode, uint32_t n, float step,
                        int32_t use_integral, float* out_rgba);
/* Build the N x N pre-integration table on the device (exact variant, or the
 * integral approximation when use_integral != 0) from the TF keys and attach
 * it to tf, replacing any table set before. */
int hipdvr_tf_preint_device(hipdvr_tf* tf, const hipdvr_tf_key* keys, int32_t nkeys, float gamma,
                            float thr_lo, float thr_hi, int32_t alpha_mode, uint32_t n, float step,
                            int32_t use_integral);

/* ---- rendering ---- */
/* Host buffers (W*H RGBA float, row 0 = bottom), synchronous. */
int hipdvr_render(hipdvr_ctx* ctx, const hipdvr_volume* vol, const hipdvr_tf* tf,
                  const float* entry_rgba, const float* exit_rgba, uint32_t w, uint32_t h,
                  const hipdvr_params* prm, float* out_rgba, hipdvr_stats* stats);
/* Device buffers, asynchronous on the context stream; stats (if not NULL)
 * are filled after an implicit synchronize. */
int hipdvr_render_device(hipdvr_ctx* ctx, const hipdvr_volume* vol, const hipdvr_tf* tf,
                         const float* d_entry, const float* d_exit, uint32_t w, uint32_t h,
                         const hipdvr_params* prm, float* d_out, hipdvr_stats* stats);
/* Entry/exit points of the volume's unit-cube proxy into device buffers. */
int hipdvr_entry_exit_device(hipdvr_ctx* ctx, const hipdvr_camera* cam, const float llf[3],
                             const float urb[3], uint32_t w, uint32_t h, int32_t quantize16,
                             float* d_entry, float* d_exit);

/* ---- kernel timing ----
 * Between hipdvr_timing_begin and hipdvr_timing_read every march launch on
 * the context is bracketed by HIP events on the context stream; read
 * synchronizes and returns the per-launch durations in milliseconds.  (The
 * reference times with a clock() based PROFILING_BLOCK,
 * include/voreen/core/processors/profiling.h:43-144.) */
int hipdvr_timing_begin(hipdvr_ctx* ctx, uint32_t capacity);
int hipdvr_timing_read(hipdvr_ctx* ctx, float* out_ms, uint32_t max, uint32_t* n);

/* ---- device buffers ---- */
int hipdvr_buffer_alloc(hipdvr_ctx* ctx, size_t bytes, void** d_ptr);
int hipdvr_buffer_free(hipdvr_ctx* ctx, void* d_ptr);
int hipdvr_memcpy_h2d(hipdvr_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int hipdvr_memcpy_d2h(hipdvr_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
/* device-to-device copy by a plain streaming kernel (16-byte vectors);
 * bytes must be a multiple of 16; asynchronous on the context stream */
int hipdvr_copy_device(hipdvr_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
