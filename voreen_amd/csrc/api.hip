// api.hip -- the C-ABI of libhipdvr.so (declared in include/hipdvr.h).
//
// Owns device objects behind opaque handles: a context (device + stream +
// events + counters), volumes (linear and bricked copies in HBM), transfer
// functions (RGBA8 table + optional pre-integration table).  No C++ exception
// crosses the ABI: every entry point catches and returns a status, and the
// message is kept per context.
#include <hip/hip_runtime.h>

#include <atomic>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hipdvr.h"
#include "ctx.h"
#include "errors.h"
#include "kernels.h"
#include "objects.h"
#include "synth_scratch.h"
#include "volume.h"

using namespace hipdvr;

namespace {

template <typename F>
int guarded(hipdvr_ctx* ctx, F&& f) {
    return ctx_guarded(ctx, std::forward<F>(f));
}

void set_device(hipdvr_ctx* ctx) { hip_check(hipSetDevice(ctx->device), "hipSetDevice"); }

uint64_t next_serial() {
    static std::atomic<uint64_t> s{1};
    return s.fetch_add(1);
}

uint64_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Grow the context's staging of hipdvr_render to `bytes` device and pinned
// host bytes (kept between frames; only a larger frame reallocates).
void ensure_stage(hipdvr_ctx* ctx, size_t bytes) {
    if (ctx->stage_bytes >= bytes) return;
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    ctx->d_stage = nullptr;
    ctx->h_stage = nullptr;
    ctx->stage_bytes = 0;
    hip_check(hipMalloc((void**)&ctx->d_stage, bytes), "hipMalloc(render staging)");
    hip_check(hipHostMalloc((void**)&ctx->h_stage, bytes, hipHostMallocDefault), "hipHostMalloc(render staging)");
    ctx->stage_bytes = bytes;
}

// memcpy of a frame between pageable and pinned host memory, split over up
// to 8 host threads in 64-byte aligned pieces (one thread below 1 MiB).
void host_copy(void* dst, const void* src, size_t bytes) {
    const size_t kMin = 1u << 20;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 8 ? 8 : nt);
    if (bytes < kMin || nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t piece = ((bytes + nt - 1) / nt + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) {
        const size_t b = piece * t;
        if (b >= bytes) break;
        const size_t n = bytes - b < piece ? bytes - b : piece;
        th.emplace_back([=] { std::memcpy((char*)dst + b, (const char*)src + b, n); });
    }
    std::memcpy(dst, src, piece < bytes ? piece : bytes);
    for (auto& x : th) x.join();
}

}  // namespace

namespace hipdvr {

// Per-brick occupancy of `vol` under the TF mapping of `tf` (and the volume's
// real-world mapping); recomputed only when the pair changes.
void ensure_occupancy(hipdvr_ctx* ctx, hipdvr_volume* vol, const hipdvr_tf* tf) {
    const uint64_t key[4] = {tf->serial, (fbits(tf->domain_lo) << 32) | fbits(tf->domain_hi),
                             (fbits(vol->desc.rwm_scale) << 32) | fbits(vol->desc.rwm_offset),
                             tf->width};
    if (std::memcmp(key, vol->occ_key, sizeof key) == 0) return;
    OccArgs o;
    o.minmax = vol->d_minmax;
    o.occ = vol->d_occ;
    o.count = vol->brick_count;
    o.dtype = vol->desc.dtype;
    o.rwm_scale = vol->desc.rwm_scale;
    o.rwm_offset = vol->desc.rwm_offset;
    o.dom_lo = tf->domain_lo;
    o.dom_hi = tf->domain_hi;
    o.tf_width = (int32_t)tf->width;
    o.alpha_prefix = tf->d_alpha_prefix;
    o.bits = vol->d_occ_bits;
    launch_occupancy(o, ctx->stream);
    hip_check(hipGetLastError(), "occupancy kernel launch");
    std::memcpy(vol->occ_key, key, sizeof key);
}

// Opacity correction per alpha byte: cpuraycaster.cpp:291 with the TF colour
// byte/255.f of :330, in float with the float overload of pow.
void fill_alpha_table(float step, float* out) {
    const float e = step * 200.0f;
    for (int k = 0; k < 256; ++k) {
        float a = (float)k / 255.f;
        out[k] = 1.f - powf(1.f - a, e);
    }
}

void validate_params(const hipdvr_params* p, const hipdvr_tf* tf) {
    if (!p) fail(HIPDVR_ERR_INVALID, "params is NULL");
    if (!(p->step > 0.f) || !std::isfinite(p->step)) fail(HIPDVR_ERR_INVALID, "step must be > 0");
    if (p->filter != HIPDVR_FILTER_NEAREST && p->filter != HIPDVR_FILTER_LINEAR)
        fail(HIPDVR_ERR_INVALID, "unknown filter mode");
    if (p->classification != HIPDVR_CLASS_TF && p->classification != HIPDVR_CLASS_PREINT &&
        p->classification != HIPDVR_CLASS_TF2D)
        fail(HIPDVR_ERR_UNSUPPORTED, "unsupported classification mode");
    if ((p->classification == HIPDVR_CLASS_TF2D) != tf->is2d)
        fail(HIPDVR_ERR_INVALID, tf->is2d ? "a 2D transfer function needs HIPDVR_CLASS_TF2D"
                                          : "HIPDVR_CLASS_TF2D needs a 2D transfer function (hipdvr_tf2d_create)");
    if (p->classification == HIPDVR_CLASS_PREINT && (!tf->d_preint || tf->preint_n < 2))
        fail(HIPDVR_ERR_INVALID, "pre-integrated classification needs a table (hipdvr_tf_set_preint)");
    if (p->gradient < HIPDVR_GRADIENT_CENTRAL || p->gradient > HIPDVR_GRADIENT_SOBEL)
        fail(HIPDVR_ERR_UNSUPPORTED, "unknown gradient mode");
    if (p->compositing < HIPDVR_COMPOSITE_DVR || p->compositing > HIPDVR_COMPOSITE_FHN)
        fail(HIPDVR_ERR_UNSUPPORTED, "unknown compositing mode");
    if (p->compositing != HIPDVR_COMPOSITE_DVR && p->classification != HIPDVR_CLASS_TF)
        fail(HIPDVR_ERR_UNSUPPORTED, "compositing modes other than DVR need 1D transfer-function classification");
    if (!std::isfinite(p->gamma) || !std::isfinite(p->iso_value))
        fail(HIPDVR_ERR_INVALID, "gamma and iso_value must be finite");
    if (p->shading < HIPDVR_SHADE_NONE || p->shading > HIPDVR_SHADE_WARD)
        fail(HIPDVR_ERR_UNSUPPORTED, "unsupported shading mode");
    if (p->kernel < HIPDVR_KERNEL_AUTO || p->kernel > HIPDVR_KERNEL_SLAB)
        fail(HIPDVR_ERR_INVALID, "unknown kernel selector");
    if (p->batch != 0 && p->batch != 1 && p->batch != 2 && p->batch != 4 && p->batch != 8)
        fail(HIPDVR_ERR_INVALID, "batch must be 0, 1, 2, 4 or 8");
}

void fill_march_args(MarchArgs& a, const hipdvr_volume* vol, const hipdvr_tf* tf,
                     const hipdvr_params* p, const float* d_entry, const float* d_exit,
                     float* d_out, uint32_t w, uint32_t h, unsigned long long* counters) {
    std::memset(&a, 0, sizeof(a));
    const VolumeDesc& v = vol->desc;
    a.vox = vol->d_linear;
    a.bricks = vol->d_tiled;
    a.cells = vol->d_cells;
    a.dtype = v.dtype;
    a.brick_occ = vol->d_occ;
    a.occ_bits = vol->d_occ_bits;
    a.occ_words = vol->occ_words;
    a.brick_log2 = v.brick_log2;
    for (int i = 0; i < 3; ++i) {
        a.dims[i] = v.dims[i];
        a.nbricks[i] = vol->nbricks[i];
        a.fdims[i] = (float)v.dims[i];
        a.fdm1[i] = (float)v.dims[i] - 1.0f;
        a.spacing[i] = v.spacing[i];
        a.offset[i] = v.offset[i];
        a.grad_off[i] = 1.f / (float)v.dims[i];
        a.grad_scale[i] = (1.f / v.spacing[i]) * 0.5f;
        a.spacing_rcp[i] = 1.f / v.spacing[i];
        a.light_pos[i] = p->light_pos[i];
        a.light_amb[i] = p->light_ambient[i];
        a.light_dif[i] = p->light_diffuse[i];
        a.light_spe[i] = p->light_specular[i];
        a.cam_pos[i] = p->camera_pos[i];
    }
    a.shininess = p->shininess;
    a.rwm_scale = v.rwm_scale;
    {
        const float range = v.dtype == HIPDVR_U8 ? 255.0f : (v.dtype == HIPDVR_U16 ? 65535.0f : 1.0f);
        for (int i = 0; i < 3; ++i) a.grad_raw[i] = (a.grad_scale[i] * a.rwm_scale) / range;
    }
    a.rwm_offset = v.rwm_offset;
    a.dom_lo = tf->domain_lo;
    a.dom_hi = tf->domain_hi;
    a.dom_unit = (tf->domain_hi - tf->domain_lo) == 1.0f ? 1 : 0;
    a.shade_model = p->shading;
    a.tf_identity = (v.rwm_scale == 1.0f && v.rwm_offset == 0.0f && tf->domain_lo == 0.0f && tf->domain_hi == 1.0f)
                        ? 1 : 0;
    a.tf_width = (int32_t)tf->width;
    a.tf = tf->d_rgba;
    a.preint = tf->d_preint;
    a.preint_n = (int32_t)tf->preint_n;
    a.step = p->step;
    a.loop_cap = p->loop_cap > 0 ? p->loop_cap : kDefaultLoopCap;
    a.entry = (const float4*)d_entry;
    a.exit = (const float4*)d_exit;
    a.out = (float4*)d_out;
    a.w = (int32_t)w;
    a.h = (int32_t)h;
    a.grid_w = (int32_t)w;
    a.grid_h = (int32_t)h;
    a.stride = 1;
    a.pix_offset = 0;
    a.cond_flag = nullptr;
    a.counters = counters;
    a.compositing = p->compositing;
    a.gamma = p->gamma;
    a.iso_value = p->iso_value;
    a.gradient = p->gradient;
    a.gvol = vol->d_grad;
    a.gvol_dtype = vol->grad_dtype;
    a.gvol_channels = (int32_t)vol->grad_channels;
    a.tf2d = tf->d_tf2d;
    a.tf2d_w = (int32_t)tf->w2d;
    a.tf2d_h = (int32_t)tf->h2d;
    for (int i = 0; i < 4; ++i) a.tf2d_dom[i] = tf->dom2d[i];
    fill_alpha_table(p->step, a.alpha_corr);
}

void ensure_tf2d(hipdvr_tf* tf, float step, hipStream_t st) {
    if (tf->d_tf2d && tf->tf2d_step == step) return;
    const size_t n = (size_t)tf->w2d * tf->h2d;
    // kernels of earlier renders may still read the device table and the
    // upload below reads the host copy asynchronously
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!tf->d_tf2d) hip_check(hipMalloc((void**)&tf->d_tf2d, n * sizeof(float4)), "hipMalloc(2D TF)");
    tf->host2d_corr = tf->host2d;
    const float e = step * 200.0f;   // samplingStepSize * SAMPLING_BASE_INTERVAL_RCP
    for (size_t i = 0; i < n; ++i) {
        float& a = tf->host2d_corr[4 * i + 3];
        a = 1.f - std::pow(1.f - a, e);
    }
    hip_check(hipMemcpyAsync(tf->d_tf2d, tf->host2d_corr.data(), n * sizeof(float4), hipMemcpyHostToDevice, st),
              "2D TF upload");
    tf->tf2d_step = step;
}

}  // namespace hipdvr

namespace {

void render_device_impl(hipdvr_ctx* ctx, const hipdvr_volume* vol, const hipdvr_tf* tf,
                        const float* d_entry, const float* d_exit, uint32_t w, uint32_t h,
                        const hipdvr_params* p, float* d_out, hipdvr_stats* stats) {
    if (!vol || !tf || !d_entry || !d_exit || !d_out) fail(HIPDVR_ERR_INVALID, "NULL argument");
    if (vol->ctx != ctx || tf->ctx != ctx) fail(HIPDVR_ERR_INVALID, "handle from another context");
    if (w == 0 || h == 0) fail(HIPDVR_ERR_INVALID, "empty image");
    validate_params(p, tf);
    set_device(ctx);
    const bool count = stats != nullptr;
    if (count) launch_zero_u64(ctx->d_counters, kNumCounters, ctx->stream);
    MarchArgs a;
    fill_march_args(a, vol, tf, p, d_entry, d_exit, d_out, w, h, ctx->d_counters);
    // any shading model: the gradient plus lighting; the fast cell variant
    // (interior gradient, hardware lighting terms) only for Phong
    const bool phong = p->shading != HIPDVR_SHADE_NONE;
    // a gradient other than central differences (shading or FHN) runs on the reference kernels
    const bool other_gradient = p->gradient != HIPDVR_GRADIENT_CENTRAL &&
                                (phong || p->compositing == HIPDVR_COMPOSITE_FHN);
    int kernel = p->kernel;
    // the skipping kernels drop samples, which the pre-integrated mode cannot
    // (its lookup pairs every sample with the previous one); the cell records
    // serve the linear filter
    const bool modes = p->compositing != HIPDVR_COMPOSITE_DVR;
    if (modes && kernel != HIPDVR_KERNEL_AUTO && kernel != HIPDVR_KERNEL_REFERENCE)
        fail(HIPDVR_ERR_UNSUPPORTED, "compositing modes other than DVR run on the reference kernel");
    if (other_gradient && kernel != HIPDVR_KERNEL_AUTO && kernel != HIPDVR_KERNEL_REFERENCE)
        fail(HIPDVR_ERR_UNSUPPORTED, "gradient modes other than central differences run on the reference kernel");
    if (p->classification == HIPDVR_CLASS_TF2D && (!vol->d_grad || vol->grad_channels < 3))
        fail(HIPDVR_ERR_INVALID, "2D transfer functions need an RGB or RGBA gradient volume "
                                 "(hipdvr_volume_set_gradient)");
    if (kernel == HIPDVR_KERNEL_AUTO) {
        if (p->classification != HIPDVR_CLASS_TF || modes || other_gradient) kernel = HIPDVR_KERNEL_REFERENCE;
        else if (p->filter == HIPDVR_FILTER_LINEAR && cells_allowed(vol)) kernel = HIPDVR_KERNEL_CELL;
        else if (p->shading == HIPDVR_SHADE_NONE || p->shading == HIPDVR_SHADE_PHONG) kernel = HIPDVR_KERNEL_BRICKED;
        else kernel = HIPDVR_KERNEL_REFERENCE;
    }
    const bool cell = kernel == HIPDVR_KERNEL_CELL;
    const bool slab = kernel == HIPDVR_KERNEL_SLAB;
    if ((kernel == HIPDVR_KERNEL_BRICKED || cell || slab) && p->classification != HIPDVR_CLASS_TF)
        fail(HIPDVR_ERR_UNSUPPORTED, "the skipping kernels support transfer-function classification only");
    if ((cell || slab) && p->filter != HIPDVR_FILTER_LINEAR)
        fail(HIPDVR_ERR_UNSUPPORTED, "the cell and slab kernels support the linear filter only");
    if (kernel == HIPDVR_KERNEL_BRICKED && p->shading != HIPDVR_SHADE_NONE && p->shading != HIPDVR_SHADE_PHONG)
        fail(HIPDVR_ERR_UNSUPPORTED, "the bricked kernel shades with Phong only");
    hipdvr_volume* mvol = const_cast<hipdvr_volume*>(vol);
    if (p->classification == HIPDVR_CLASS_TF2D) {
        hipdvr_tf* mtf = const_cast<hipdvr_tf*>(tf);
        ensure_tf2d(mtf, p->step, ctx->stream);
        a.tf2d = mtf->d_tf2d;
    }
    if (kernel == HIPDVR_KERNEL_BRICKED || cell || slab) ensure_occupancy(ctx, mvol, tf);
    if (cell) {
        ensure_cells(mvol, ctx->stream);
        a.cells = mvol->d_cells;
    }
    if (count) hip_check(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
    const bool timed = ctx_timing_start(ctx);
    if (modes) {
        launch_march_modes(a, vol->desc.dtype, p->filter, phong, ctx->stream);
    } else if (kernel == HIPDVR_KERNEL_REFERENCE) {
        launch_march_ref(a, vol->desc.dtype, p->filter, p->classification, phong, count, ctx->stream);
    } else if (slab) {
        const int shade = phong ? ((p->exact_shading || p->shading != HIPDVR_SHADE_PHONG) ? 2 : 1) : 0;
        launch_march_slab(a, vol->desc.dtype, shade, count, ctx->stream);
    } else if (kernel == HIPDVR_KERNEL_CELL) {
        // samples in flight per ray: default_batch (kernels.h)
        const int shade = phong ? ((p->exact_shading || p->shading != HIPDVR_SHADE_PHONG) ? 2 : 1) : 0;
        launch_march_cell(a, vol->desc.dtype, shade, count, mvol->cells_wide,
                          p->batch ? p->batch : default_batch(shade), ctx->stream);

    } else {
        // buffer descriptors address at most 2^32 - 1 bytes
        const uint64_t vsz = vol->desc.dtype == HIPDVR_U8 ? 1 : 2;
        const bool wide = vol->tiled_elems * vsz > 0xffffff00ull;
        launch_march_fast(a, vol->desc.dtype, p->filter, phong, count, wide, ctx->stream);
    }
    hip_check(hipGetLastError(), "march kernel launch");
    if (timed) ctx_timing_stop(ctx);
    if (count) {
        hip_check(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord");
        unsigned long long c[kNumCounters];
        hip_check(hipMemcpyAsync(c, ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, ctx->stream),
                  "hipMemcpyAsync(counters)");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
        stats->rays = c[0];
        stats->samples = c[1];
        stats->fetched = c[2];
        stats->shaded = c[3];
        stats->lane_slots = c[4];
        stats->shade_slots = c[5];
        stats->kernel_ms = ms;
    }
}

// Camera -> texture-space ray-generation constants.  Double precision, then
// rounded to float; the per-pixel part runs in float in the kernel.
RayGen raygen_setup(const hipdvr_camera* cam, const float llf[3], const float urb[3], uint32_t w,
                    uint32_t h, int quantize) {
    double f[3], up[3], r[3], u[3], n;
    for (int i = 0; i < 3; ++i) {
        f[i] = (double)cam->focus[i] - cam->eye[i];
        up[i] = cam->up[i];
    }
    n = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int i = 0; i < 3; ++i) f[i] /= n;
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int i = 0; i < 3; ++i) r[i] /= n;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    // tgt::Camera::getProjectionMatrix (ext/tgt/camera.cpp:203-214), frustum
    // ratio 1: fovy, clamped to [6, 175] degrees, spans the image height when
    // the window is at least as wide as high, else its width
    const double fovy = cam->fovy_deg < 6.0f ? 6.0 : (cam->fovy_deg > 175.0f ? 175.0 : (double)cam->fovy_deg);
    const double th = std::tan(fovy * 0.5 * 3.14159265358979323846 / 180.0);
    const double aspect = (double)w / (double)h;
    const double th_u = aspect >= 1.0 ? th : th / aspect;   // half height at unit view distance
    const double th_r = th_u * aspect;                      // half width
    RayGen g;
    for (int i = 0; i < 3; ++i) {
        const double inv = 1.0 / ((double)urb[i] - (double)llf[i]);
        g.o[i] = (float)(((double)cam->eye[i] - (double)llf[i]) * inv);
        g.f[i] = (float)(f[i] * inv);
        g.r[i] = (float)(r[i] * th_r * inv);
        g.u[i] = (float)(u[i] * th_u * inv);
    }
    g.near_t = cam->near_dist;
    g.w = (int32_t)w;
    g.h = (int32_t)h;
    g.quantize = quantize;
    return g;
}

}  // namespace

extern "C" {

int hipdvr_abi_version(void) { return HIPDVR_ABI_VERSION; }

const char* hipdvr_last_error(const hipdvr_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "NULL context";
}

int hipdvr_create(hipdvr_ctx** out, int device) {
    if (!out) return HIPDVR_ERR_INVALID;
    *out = nullptr;
    hipdvr_ctx* ctx = new (std::nothrow) hipdvr_ctx();
    if (!ctx) return HIPDVR_ERR_NOMEM;
    int rc = guarded(ctx, [&] {
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) fail(HIPDVR_ERR_INVALID, "device index out of range");
        ctx->device = device;
        set_device(ctx);
        hip_check(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking), "hipStreamCreate");
        ctx->stream = ctx->own_stream;
        hip_check(hipEventCreate(&ctx->ev0), "hipEventCreate");
        hip_check(hipEventCreate(&ctx->ev1), "hipEventCreate");
        hip_check(hipMalloc((void**)&ctx->d_counters, kNumCounters * sizeof(unsigned long long)), "hipMalloc(counters)");
        // the no-allocation generator's scratch address on this device
        synth_scratch_init(device);
    });
    if (rc != HIPDVR_OK) {
        std::fprintf(stderr, "hipdvr_create: %s\n", ctx->last_error.c_str());
        hipdvr_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return HIPDVR_OK;
}

int hipdvr_destroy(hipdvr_ctx* ctx) {
    if (!ctx) return HIPDVR_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    if (ctx->stream && ctx->stream != ctx->own_stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->tev) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return HIPDVR_OK;
}

int hipdvr_set_stream(hipdvr_ctx* ctx, void* stream) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return HIPDVR_OK;
}

int hipdvr_synchronize(hipdvr_ctx* ctx) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        set_device(ctx);
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int hipdvr_volume_create(hipdvr_ctx* ctx, const hipdvr_volume_desc* desc, const void* host_voxels,
                         hipdvr_volume** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!out || !desc || !host_voxels) fail(HIPDVR_ERR_INVALID, "NULL argument");
        set_device(ctx);
        *out = volume_create(ctx, ctx->stream, desc, host_voxels, /*device=*/false);
    });
}

int hipdvr_volume_create_device(hipdvr_ctx* ctx, const hipdvr_volume_desc* desc,
                                const void* device_voxels, hipdvr_volume** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!out || !desc || !device_voxels) fail(HIPDVR_ERR_INVALID, "NULL argument");
        set_device(ctx);
        *out = volume_create(ctx, ctx->stream, desc, device_voxels, /*device=*/true);
    });
}

int hipdvr_volume_destroy(hipdvr_volume* vol) {
    if (!vol) return HIPDVR_OK;
    volume_destroy(vol);
    return HIPDVR_OK;
}

int hipdvr_volume_device_ptr(const hipdvr_volume* vol, const void** out) {
    if (!vol || !out) return HIPDVR_ERR_INVALID;
    *out = vol->d_linear;
    return HIPDVR_OK;
}

int hipdvr_tf_create(hipdvr_ctx* ctx, const uint8_t* rgba8, uint32_t width, float domain_lo,
                     float domain_hi, hipdvr_tf** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!out || !rgba8) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (width < 2 || width > (uint32_t)kMaxTfWidth)
            fail(HIPDVR_ERR_INVALID, "TF width must be in [2, 4096]");
        if (!(domain_lo <= domain_hi)) fail(HIPDVR_ERR_INVALID, "invalid transfer function domain");
        set_device(ctx);
        hipdvr_tf* tf = new hipdvr_tf();
        tf->ctx = ctx;
        tf->width = width;
        tf->domain_lo = domain_lo;
        tf->domain_hi = domain_hi;
        tf->serial = next_serial();
        tf->host_rgba.assign(rgba8, rgba8 + 4 * (size_t)width);
        std::vector<uint32_t> prefix(width + 1, 0);
        for (uint32_t i = 0; i < width; ++i) prefix[i + 1] = prefix[i] + (rgba8[4 * i + 3] != 0);
        hipError_t e = hipMalloc((void**)&tf->d_rgba, 4 * (size_t)width);
        if (e == hipSuccess)
            e = hipMemcpy(tf->d_rgba, rgba8, 4 * (size_t)width, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void**)&tf->d_alpha_prefix, 4 * (size_t)(width + 1));
        if (e == hipSuccess)
            e = hipMemcpy(tf->d_alpha_prefix, prefix.data(), 4 * (size_t)(width + 1), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (tf->d_rgba) (void)hipFree(tf->d_rgba);
            if (tf->d_alpha_prefix) (void)hipFree(tf->d_alpha_prefix);
            delete tf;
            hip_check(e, "TF upload");
        }
        *out = tf;
    });
}

int hipdvr_tf_destroy(hipdvr_tf* tf) {
    if (!tf) return HIPDVR_OK;
    (void)hipSetDevice(tf->ctx->device);
    if (tf->d_rgba) (void)hipFree(tf->d_rgba);
    if (tf->d_alpha_prefix) (void)hipFree(tf->d_alpha_prefix);
    if (tf->d_preint) (void)hipFree(tf->d_preint);
    if (tf->d_tf2d) (void)hipFree(tf->d_tf2d);
    delete tf;
    return HIPDVR_OK;
}

int hipdvr_tf2d_create(hipdvr_ctx* ctx, const float* rgba, uint32_t width, uint32_t height,
                       const float domain[4], hipdvr_tf** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!out || !rgba || !domain) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (width < 2 || height < 2 || width > 4096 || height > 4096)
            fail(HIPDVR_ERR_INVALID, "2D TF width and height must be in [2, 4096]");
        if (!(domain[0] <= domain[1]) || !(domain[2] <= domain[3]))
            fail(HIPDVR_ERR_INVALID, "invalid 2D transfer function domain");
        hipdvr_tf* tf = new hipdvr_tf();
        tf->ctx = ctx;
        tf->serial = next_serial();
        tf->is2d = true;
        tf->w2d = width;
        tf->h2d = height;
        for (int i = 0; i < 4; ++i) tf->dom2d[i] = domain[i];
        tf->host2d.assign(rgba, rgba + 4 * (size_t)width * height);
        *out = tf;
    });
}

int hipdvr_volume_set_gradient(hipdvr_volume* vol, const void* host_data, int32_t dtype, uint32_t channels) {
    if (!vol) return HIPDVR_ERR_INVALID;
    return guarded(vol->ctx, [&] {
        if (!host_data) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (dtype != HIPDVR_U8 && dtype != HIPDVR_U16 && dtype != HIPDVR_F32)
            fail(HIPDVR_ERR_INVALID, "gradient volume dtype must be HIPDVR_U8, HIPDVR_U16 or HIPDVR_F32");
        if (channels != 3 && channels != 4)
            fail(HIPDVR_ERR_INVALID, "gradient volumes have 3 (RGB) or 4 (RGBA) channels");
        const size_t esz = dtype == HIPDVR_U8 ? 1 : (dtype == HIPDVR_U16 ? 2 : 4);
        const size_t bytes = (size_t)vol->desc.dims[0] * vol->desc.dims[1] * vol->desc.dims[2] * channels * esz;
        set_device(vol->ctx);
        hip_check(hipStreamSynchronize(vol->ctx->stream), "hipStreamSynchronize");
        void* d = nullptr;
        hip_check(hipMalloc(&d, bytes), "hipMalloc(gradient volume)");
        const hipError_t e = hipMemcpy(d, host_data, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            hip_check(e, "gradient volume upload");
        }
        if (vol->d_grad) (void)hipFree(vol->d_grad);
        vol->d_grad = d;
        vol->grad_dtype = dtype;
        vol->grad_channels = channels;
    });
}

int hipdvr_tf_set_preint(hipdvr_tf* tf, const float* table, uint32_t n) {
    if (!tf) return HIPDVR_ERR_INVALID;
    return guarded(tf->ctx, [&] {
        if (!table || n < 2 || n > 4096) fail(HIPDVR_ERR_INVALID, "pre-integration table size must be in [2, 4096]");
        set_device(tf->ctx);
        if (tf->d_preint) {
            (void)hipFree(tf->d_preint);
            tf->d_preint = nullptr;
            tf->preint_n = 0;
        }
        const size_t bytes = (size_t)n * n * sizeof(float4);
        hip_check(hipMalloc((void**)&tf->d_preint, bytes), "hipMalloc(preint)");
        hip_check(hipMemcpy(tf->d_preint, table, bytes, hipMemcpyHostToDevice), "preint upload");
        tf->preint_n = n;
    });
}

int hipdvr_render_device(hipdvr_ctx* ctx, const hipdvr_volume* vol, const hipdvr_tf* tf,
                         const float* d_entry, const float* d_exit, uint32_t w, uint32_t h,
                         const hipdvr_params* prm, float* d_out, hipdvr_stats* stats) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] { render_device_impl(ctx, vol, tf, d_entry, d_exit, w, h, prm, d_out, stats); });
}

int hipdvr_render(hipdvr_ctx* ctx, const hipdvr_volume* vol, const hipdvr_tf* tf,
                  const float* entry, const float* exit_, uint32_t w, uint32_t h,
                  const hipdvr_params* prm, float* out, hipdvr_stats* stats) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!entry || !exit_ || !out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (w == 0 || h == 0) fail(HIPDVR_ERR_INVALID, "empty image");
        set_device(ctx);
        const size_t px = (size_t)w * h * 4, bytes = px * sizeof(float);
        ensure_stage(ctx, 3 * bytes);
        float *de = ctx->d_stage, *dx = de + px, *dout = dx + px;
        float *pe = ctx->h_stage, *pxb = pe + px, *po = pxb + px;
        // the caller's pageable buffers through the pinned ones: each copy
        // into pinned memory is split over host threads, and the entry
        // upload runs while the exit points are copied
        host_copy(pe, entry, bytes);
        hip_check(hipMemcpyAsync(de, pe, bytes, hipMemcpyHostToDevice, ctx->stream), "upload entry");
        host_copy(pxb, exit_, bytes);
        hip_check(hipMemcpyAsync(dx, pxb, bytes, hipMemcpyHostToDevice, ctx->stream), "upload exit");
        render_device_impl(ctx, vol, tf, de, dx, w, h, prm, dout, stats);
        hip_check(hipMemcpyAsync(po, dout, bytes, hipMemcpyDeviceToHost, ctx->stream), "download out");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        host_copy(out, po, bytes);
    });
}

int hipdvr_entry_exit_device(hipdvr_ctx* ctx, const hipdvr_camera* cam, const float llf[3],
                             const float urb[3], uint32_t w, uint32_t h, int32_t quantize16,
                             float* d_entry, float* d_exit) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!cam || !llf || !urb || !d_entry || !d_exit) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (w == 0 || h == 0) fail(HIPDVR_ERR_INVALID, "empty image");
        for (int i = 0; i < 3; ++i)
            if (!(urb[i] > llf[i])) fail(HIPDVR_ERR_INVALID, "proxy box must have urb > llf");
        set_device(ctx);
        RayGen g = raygen_setup(cam, llf, urb, w, h, quantize16 ? 1 : 0);
        launch_entry_exit(g, (float4*)d_entry, (float4*)d_exit, ctx->stream);
        hip_check(hipGetLastError(), "entry/exit kernel launch");
    });
}

int hipdvr_timing_begin(hipdvr_ctx* ctx, uint32_t capacity) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (capacity > (1u << 16)) fail(HIPDVR_ERR_INVALID, "timing capacity too large");
        set_device(ctx);
        while (ctx->tev.size() < 2 * (size_t)capacity) {
            hipEvent_t e;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            ctx->tev.push_back(e);
        }
        ctx->tcap = capacity;
        ctx->tcount = 0;
    });
}

int hipdvr_timing_read(hipdvr_ctx* ctx, float* out_ms, uint32_t max, uint32_t* n) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!n || (max && !out_ms)) fail(HIPDVR_ERR_INVALID, "NULL argument");
        set_device(ctx);
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        const uint32_t k = ctx->tcount < max ? ctx->tcount : max;
        for (uint32_t i = 0; i < k; ++i)
            hip_check(hipEventElapsedTime(&out_ms[i], ctx->tev[2 * i], ctx->tev[2 * i + 1]),
                      "hipEventElapsedTime");
        *n = ctx->tcount;
        ctx->tcap = 0;
        ctx->tcount = 0;
    });
}

int hipdvr_copy_device(hipdvr_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!d_dst || !d_src || bytes % 16) fail(HIPDVR_ERR_INVALID, "copy needs 16-byte multiples");
        set_device(ctx);
        launch_copy16(d_dst, d_src, bytes / 16, ctx->stream);
        hip_check(hipGetLastError(), "copy kernel launch");
    });
}

int hipdvr_buffer_alloc(hipdvr_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        if (!d_ptr || bytes == 0) fail(HIPDVR_ERR_INVALID, "bad buffer request");
        set_device(ctx);
        *d_ptr = nullptr;
        hip_check(hipMalloc(d_ptr, bytes), "hipMalloc");
    });
}

int hipdvr_buffer_free(hipdvr_ctx* ctx, void* d_ptr) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        set_device(ctx);
        if (d_ptr) hip_check(hipFree(d_ptr), "hipFree");
    });
}

int hipdvr_memcpy_h2d(hipdvr_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        set_device(ctx);
        hip_check(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int hipdvr_memcpy_d2h(hipdvr_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return guarded(ctx, [&] {
        set_device(ctx);
        hip_check(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

}  // extern "C"
