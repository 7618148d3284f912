// learned.hip -- C-ABI of the stepping network (include/hipdvr_learned.h):
// creation, parameter I/O, forward, gradient and Adam training steps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hipdvr_learned.h"
#include "ctx.h"
#include "errors.h"
#include "kernels.h"
#include "learned.h"
#include "mlp.h"
#include "objects.h"

using namespace hipdvr;

struct hipdvr_mlp {
    hipdvr_ctx* ctx = nullptr;
    float* d_master = nullptr;   // FP32 parameters
    float* d_m = nullptr;        // Adam moments
    float* d_v = nullptr;
    float* d_grad = nullptr;
    float* d_loss = nullptr;
    MlpDevParams dev{};
    int step = 0;
};

static_assert(kMlpParams == HIPDVR_MLP_PARAMS, "parameter count mismatch");

namespace {

void free_mlp(hipdvr_mlp* m) {
    if (!m) return;
    (void)hipFree(m->d_master);
    (void)hipFree(m->d_m);
    (void)hipFree(m->d_v);
    (void)hipFree(m->d_grad);
    (void)hipFree(m->d_loss);
    (void)hipFree(m->dev.w1);
    (void)hipFree(m->dev.w2);
    (void)hipFree(m->dev.w3);
    (void)hipFree(m->dev.b);
    delete m;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int hipdvr_mlp_create(hipdvr_ctx* ctx, uint64_t seed, hipdvr_mlp** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        if (!out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        ctx_set_device(ctx);
        hipdvr_mlp* m = new hipdvr_mlp();
        m->ctx = ctx;
        try {
            const size_t pb = kMlpParams * sizeof(float);
            hip_check(hipMalloc((void**)&m->d_master, pb), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->d_m, pb), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->d_v, pb), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->d_grad, pb + 64), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->d_loss, 64), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->dev.w1, kMlpIn * kMlpHidden * 2), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->dev.w2, kMlpHidden * kMlpHidden * 2), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->dev.w3, kMlpHidden * 16 * 2), "hipMalloc(mlp)");
            hip_check(hipMalloc((void**)&m->dev.b, (2 * kMlpHidden + 16) * sizeof(float)), "hipMalloc(mlp)");
            // Xavier-uniform weights, zero biases
            std::vector<float> w(kMlpParams, 0.f);
            uint64_t s = seed;
            auto fill = [&](size_t off, int fan_in, int fan_out) {
                const double lim = std::sqrt(6.0 / (fan_in + fan_out));
                for (int i = 0; i < fan_in * fan_out; ++i) {
                    const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                    w[off + i] = (float)((2.0 * u - 1.0) * lim);
                }
            };
            fill(0, kMlpIn, kMlpHidden);
            fill(kMlpIn * kMlpHidden + kMlpHidden, kMlpHidden, kMlpHidden);
            fill(kMlpIn * kMlpHidden + 2 * kMlpHidden + kMlpHidden * kMlpHidden, kMlpHidden, kMlpOut);
            std::vector<float> zero(kMlpParams, 0.f);
            hip_check(hipMemcpy(m->d_master, w.data(), pb, hipMemcpyHostToDevice), "mlp upload");
            hip_check(hipMemcpy(m->d_m, zero.data(), pb, hipMemcpyHostToDevice), "mlp upload");
            hip_check(hipMemcpy(m->d_v, zero.data(), pb, hipMemcpyHostToDevice), "mlp upload");
            launch_mlp_pack(m->d_master, m->dev, ctx_stream(ctx));
            hip_check(hipGetLastError(), "mlp pack");
            hip_check(hipStreamSynchronize(ctx_stream(ctx)), "hipStreamSynchronize");
        } catch (...) {
            free_mlp(m);
            throw;
        }
        *out = m;
    });
}

int hipdvr_mlp_destroy(hipdvr_mlp* m) {
    if (!m) return HIPDVR_OK;
    (void)hipStreamSynchronize(ctx_stream(m->ctx));
    free_mlp(m);
    return HIPDVR_OK;
}

int hipdvr_mlp_get_params(hipdvr_mlp* m, float* out) {
    if (!m) return HIPDVR_ERR_INVALID;
    return ctx_guarded(m->ctx, [&] {
        if (!out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        ctx_set_device(m->ctx);
        hip_check(hipStreamSynchronize(ctx_stream(m->ctx)), "hipStreamSynchronize");
        hip_check(hipMemcpy(out, m->d_master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost), "mlp download");
    });
}

int hipdvr_mlp_set_params(hipdvr_mlp* m, const float* in) {
    if (!m) return HIPDVR_ERR_INVALID;
    return ctx_guarded(m->ctx, [&] {
        if (!in) fail(HIPDVR_ERR_INVALID, "NULL argument");
        for (int i = 0; i < kMlpParams; ++i)
            if (!std::isfinite(in[i])) fail(HIPDVR_ERR_INVALID, "non-finite parameter");
        ctx_set_device(m->ctx);
        hip_check(hipStreamSynchronize(ctx_stream(m->ctx)), "hipStreamSynchronize");
        hip_check(hipMemcpy(m->d_master, in, kMlpParams * sizeof(float), hipMemcpyHostToDevice), "mlp upload");
        launch_mlp_pack(m->d_master, m->dev, ctx_stream(m->ctx));
        hip_check(hipGetLastError(), "mlp pack");
        hip_check(hipStreamSynchronize(ctx_stream(m->ctx)), "hipStreamSynchronize");
    });
}

int hipdvr_mlp_forward(hipdvr_mlp* m, const float* d_x, uint32_t rows, float* d_y) {
    if (!m) return HIPDVR_ERR_INVALID;
    return ctx_guarded(m->ctx, [&] {
        if (!d_x || !d_y) fail(HIPDVR_ERR_INVALID, "NULL argument");
        ctx_set_device(m->ctx);
        launch_mlp_forward(m->dev, d_x, rows, d_y, ctx_stream(m->ctx));
        hip_check(hipGetLastError(), "mlp forward launch");
    });
}

int hipdvr_mlp_gradient(hipdvr_mlp* m, const float* d_x, const float* d_t, uint32_t rows,
                        float* grad_out, float* loss_out) {
    if (!m) return HIPDVR_ERR_INVALID;
    return ctx_guarded(m->ctx, [&] {
        if (!d_x || !d_t || !grad_out || rows == 0) fail(HIPDVR_ERR_INVALID, "bad arguments");
        ctx_set_device(m->ctx);
        launch_mlp_grad(m->dev, d_x, d_t, rows, m->d_grad, m->d_loss, ctx_stream(m->ctx));
        hip_check(hipGetLastError(), "mlp grad launch");
        hip_check(hipMemcpyAsync(grad_out, m->d_grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
                                 ctx_stream(m->ctx)), "grad download");
        float l = 0.f;
        hip_check(hipMemcpyAsync(&l, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, ctx_stream(m->ctx)),
                  "loss download");
        hip_check(hipStreamSynchronize(ctx_stream(m->ctx)), "hipStreamSynchronize");
        if (loss_out) *loss_out = l;
    });
}

int hipdvr_mlp_train_step(hipdvr_mlp* m, const float* d_x, const float* d_t, uint32_t rows, float lr,
                          float* loss_out) {
    if (!m) return HIPDVR_ERR_INVALID;
    return ctx_guarded(m->ctx, [&] {
        if (!d_x || !d_t || rows == 0 || !(lr > 0.f)) fail(HIPDVR_ERR_INVALID, "bad arguments");
        ctx_set_device(m->ctx);
        hipStream_t st = ctx_stream(m->ctx);
        launch_mlp_grad(m->dev, d_x, d_t, rows, m->d_grad, m->d_loss, st);
        m->step += 1;
        launch_mlp_adam(m->d_master, m->d_m, m->d_v, m->d_grad, lr, m->step, st);
        launch_mlp_pack(m->d_master, m->dev, st);
        hip_check(hipGetLastError(), "mlp train launch");
        if (loss_out) {
            hip_check(hipMemcpyAsync(loss_out, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, st),
                      "loss download");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
    });
}

}  // extern "C"

namespace hipdvr {
const MlpDevParams& mlp_dev_params(const hipdvr_mlp* m) { return m->dev; }
hipdvr_ctx* mlp_ctx(const hipdvr_mlp* m) { return m->ctx; }
}  // namespace hipdvr

// ---------------------------------------------------------------------------
// learned rendering
// ---------------------------------------------------------------------------
struct hipdvr_learned {
    hipdvr_ctx* ctx = nullptr;
    const hipdvr_volume* vol = nullptr;
    uint32_t nb[3] = {0, 0, 0};
    uint32_t bricks = 0, rows = 0;
    float4* d_stats = nullptr;
    uint8_t* d_occ16 = nullptr;
    float* d_feat = nullptr;         // [rows][16]
    float* d_y = nullptr;            // [rows][2]
    uint8_t* d_table = nullptr;      // [rows]
    float* d_err = nullptr;          // [rows][4]
    unsigned int* d_visits = nullptr;
    float* d_alpha = nullptr;        // [4096] corrected alpha (base step)
    float* d_alpha_prefix = nullptr; // [4097]
    uint32_t* d_nz_prefix = nullptr; // [4097]
    float* d_alpha_m = nullptr;      // [8][256] corrected alpha for step * 1 .. 8
    std::vector<float> h_alpha_m;    // host copy (source of the last upload)
    float alpha_m_step = -1.f;       // step the tables were made for
    const hipdvr_tf* prepared_tf = nullptr;   // TF of the last hipdvr_learned_prepare
    bool gains = true;               // the prepared table departs from the plain march
    float4* d_sparse = nullptr;      // sparse plain image
    size_t sparse_cap = 0;
    float4* d_seg = nullptr;         // its ray segments
    size_t seg_cap = 0;
    float* d_sse = nullptr;
    int* d_flag = nullptr;
    unsigned int* d_count = nullptr;
    // the sparse plain re-march runs beside the learned march on a second
    // stream (default priority): its few rays would leave the GPU idle alone
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace hipdvr {
hipdvr_ctx* learned_ctx(const hipdvr_learned* l) { return l->ctx; }
const hipdvr_volume* learned_volume(const hipdvr_learned* l) { return l->vol; }
const hipdvr_tf* learned_prepared_tf(const hipdvr_learned* l) { return l->prepared_tf; }
bool learned_gains(const hipdvr_learned* l) { return l->gains; }
}  // namespace hipdvr

namespace {

// The judge's sparse pixels, inside the image, and device buffers sized for
// them: along each axis blocks of kJudgeBlock consecutive pixels, one block
// in every stride * kJudgeBlock pixels, centred in it (judge_pixel).  The
// same 1/stride^2 of the pixels as a uniform grid, but neighbouring judge
// rays share their cell records in the caches: on 512^3 / 1024^2 with Phong
// the re-march takes 0.70 ms (segments 0.43, join 0.28) against 0.78 ms for
// every stride-th pixel, and the learned frame runs 1161 against 1090
// frames/s.
constexpr uint32_t kJudgeBlock = 8;

struct SparseGrid {
    uint32_t stride, block, offs;
    int w, h;
};

// Grid size and placement only (host arithmetic, no allocation).
SparseGrid sparse_dims(uint32_t w, uint32_t h, uint32_t stride) {
    SparseGrid g;
    g.stride = stride;
    g.block = kJudgeBlock;
    g.offs = (stride - 1) * kJudgeBlock / 2;
    if (g.offs > w - 1) g.offs = w - 1;
    if (g.offs > h - 1) g.offs = h - 1;
    // grid columns g with judge_pixel(g) <= n - 1
    auto count = [&](uint32_t n) {
        const uint32_t q = n - 1 - g.offs, span = stride * kJudgeBlock, r = q % span + 1;
        return (int)((q / span) * kJudgeBlock + (r < kJudgeBlock ? r : kJudgeBlock));
    };
    g.w = count(w);
    g.h = count(h);
    return g;
}

SparseGrid sparse_grid(hipdvr_learned* l, uint32_t w, uint32_t h, uint32_t stride) {
    const SparseGrid g = sparse_dims(w, h, stride);
    const size_t need = (size_t)g.w * g.h * sizeof(float4);
    if (need > l->sparse_cap) {
        if (l->d_sparse) (void)hipFree(l->d_sparse);
        l->d_sparse = nullptr;
        l->sparse_cap = 0;
        hip_check(hipMalloc((void**)&l->d_sparse, need), "hipMalloc(sparse)");
        l->sparse_cap = need;
    }
    const size_t seg = sparse_segment_bytes((uint32_t)g.w * (uint32_t)g.h);
    if (seg > l->seg_cap) {
        if (l->d_seg) (void)hipFree(l->d_seg);
        l->d_seg = nullptr;
        l->seg_cap = 0;
        hip_check(hipMalloc((void**)&l->d_seg, seg), "hipMalloc(sparse segments)");
        l->seg_cap = seg;
    }
    return g;
}

// The plain re-march of the sparse pixels into d_out (g.w x g.h RGBA).
void launch_sparse(hipdvr_learned* l, hipdvr_volume* vol, const hipdvr_tf* tf, const hipdvr_params* prm,
                   const float* d_entry, const float* d_exit, uint32_t w, uint32_t h, const SparseGrid& g, int shade,
                   float* d_out, hipStream_t st) {
    MarchArgs sp;
    fill_march_args(sp, vol, tf, prm, d_entry, d_exit, d_out, w, h, nullptr);
    sp.cells = vol->d_cells;
    sp.grid_w = g.w;
    sp.grid_h = g.h;
    sp.stride = (int32_t)g.stride;
    sp.pix_block = (int32_t)g.block;
    sp.pix_offset = (int32_t)g.offs;
    launch_sparse_remarch(sp, vol->desc.dtype, shade, vol->cells_wide, l->d_seg, st);
}

void free_learned(hipdvr_learned* l) {
    if (!l) return;
    (void)hipFree(l->d_stats);
    (void)hipFree(l->d_occ16);
    (void)hipFree(l->d_feat);
    (void)hipFree(l->d_y);
    (void)hipFree(l->d_table);
    (void)hipFree(l->d_err);
    (void)hipFree(l->d_visits);
    (void)hipFree(l->d_alpha);
    (void)hipFree(l->d_alpha_prefix);
    (void)hipFree(l->d_nz_prefix);
    (void)hipFree(l->d_alpha_m);
    (void)hipFree(l->d_sparse);
    (void)hipFree(l->d_seg);
    (void)hipFree(l->d_sse);
    (void)hipFree(l->d_flag);
    (void)hipFree(l->d_count);
    if (l->fork) (void)hipEventDestroy(l->fork);
    if (l->join) (void)hipEventDestroy(l->join);
    if (l->side) (void)hipStreamDestroy(l->side);
    delete l;
}

hipdvr_learned_opts default_opts(const hipdvr_learned_opts* o) {
    hipdvr_learned_opts d;
    d.psnr_floor = 45.f;
    d.remarch_stride = 8;
    d.skip_logit = 0.f;
    d.max_log2_step = 3;
    if (o) {
        if (o->psnr_floor > 0.f) d.psnr_floor = o->psnr_floor;
        if (o->remarch_stride > 0) d.remarch_stride = o->remarch_stride > 1024 ? 1024 : o->remarch_stride;
        d.skip_logit = o->skip_logit;
        d.max_log2_step = o->max_log2_step < 0 ? 0 : (o->max_log2_step > 3 ? 3 : o->max_log2_step);
    }
    return d;
}

// Corrected alpha per TF index at the base step, its prefix sums and the
// prefix counts of non-zero alpha bytes (feature inputs).
void upload_tf_tables(hipdvr_learned* l, const hipdvr_tf* tf, float step, hipStream_t st) {
    float corr[256];
    fill_alpha_table(step, corr);
    const uint32_t w = tf->width;
    std::vector<float> alpha(w), pre(w + 1, 0.f);
    std::vector<uint32_t> nz(w + 1, 0);
    for (uint32_t i = 0; i < w; ++i) {
        const uint8_t ab = tf->host_rgba[4 * i + 3];
        alpha[i] = corr[ab];
        pre[i + 1] = pre[i] + alpha[i];
        nz[i + 1] = nz[i] + (ab != 0);
    }
    hip_check(hipMemcpyAsync(l->d_alpha, alpha.data(), w * 4, hipMemcpyHostToDevice, st), "upload");
    hip_check(hipMemcpyAsync(l->d_alpha_prefix, pre.data(), (w + 1) * 4, hipMemcpyHostToDevice, st), "upload");
    hip_check(hipMemcpyAsync(l->d_nz_prefix, nz.data(), (w + 1) * 4, hipMemcpyHostToDevice, st), "upload");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// Opacity-correction tables for step multiples 1..8, uploaded when the step
// changes.
void ensure_alpha_tables(hipdvr_learned* l, float step, hipStream_t st) {
    if (l->alpha_m_step == step) return;
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");   // previous upload done with the host copy
    l->h_alpha_m.assign((size_t)kLMaxMult * 256, 0.f);
    for (int k = 0; k < kLMaxMult; ++k) fill_alpha_table(step * (float)(k + 1), &l->h_alpha_m[(size_t)k * 256]);
    hip_check(hipMemcpyAsync(l->d_alpha_m, l->h_alpha_m.data(), l->h_alpha_m.size() * sizeof(float),
                             hipMemcpyHostToDevice, st), "upload");
    l->alpha_m_step = step;
}

void fill_learned_args(LearnedArgs& la, hipdvr_learned* l, const hipdvr_tf* tf, const hipdvr_params* p,
                       const float* d_entry, const float* d_exit, float* d_out, uint32_t w, uint32_t h) {
    ensure_alpha_tables(l, p->step, ctx_stream(l->ctx));
    std::memset(&la, 0, sizeof(la));
    fill_march_args(la.m, l->vol, tf, p, d_entry, d_exit, d_out, w, h, l->ctx->d_counters);
    la.u16 = l->vol->desc.dtype == HIPDVR_U16 ? 1 : 0;
    for (int i = 0; i < 3; ++i) la.nb16[i] = l->nb[i];
    la.table = l->d_table;
    la.occ16 = l->d_occ16;
    la.alpha_m = l->d_alpha_m;
}

}  // namespace

namespace hipdvr {

JudgeGrid learned_judge_grid(uint32_t w, uint32_t h, const hipdvr_learned_opts& opts) {
    const SparseGrid g = sparse_dims(w, h, (uint32_t)opts.remarch_stride);
    return JudgeGrid{g.stride, g.block, g.offs, g.w, g.h};
}

hipdvr_learned_opts learned_options(const hipdvr_learned_opts* o) { return default_opts(o); }

void learned_ensure_step(hipdvr_learned* l, float step) { ensure_alpha_tables(l, step, ctx_stream(l->ctx)); }

void learned_stepping_args(hipdvr_learned* l, const hipdvr_params* p, LearnedArgs& la) {
    ensure_alpha_tables(l, p->step, ctx_stream(l->ctx));
    la.u16 = l->vol->desc.dtype == HIPDVR_U16 ? 1 : 0;
    for (int i = 0; i < 3; ++i) la.nb16[i] = l->nb[i];
    la.table = l->d_table;
    la.occ16 = l->d_occ16;
    la.alpha_m = l->d_alpha_m;
}

}  // namespace hipdvr

extern "C" {

int hipdvr_learned_create(hipdvr_ctx* ctx, const hipdvr_volume* vol, hipdvr_learned** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        if (!vol || !out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (vol->ctx != ctx) fail(HIPDVR_ERR_INVALID, "volume from another context");
        if (!cells_allowed(vol))
            fail(HIPDVR_ERR_UNSUPPORTED, "the volume's cell records exceed the library's size limit");
        ctx_set_device(ctx);
        hipdvr_learned* l = new hipdvr_learned();
        l->ctx = ctx;
        l->vol = vol;
        try {
            l->bricks = 1;
            for (int i = 0; i < 3; ++i) {
                l->nb[i] = (vol->desc.dims[i] + kLBrick - 1) / kLBrick;
                l->bricks *= l->nb[i];
            }
            l->rows = l->bricks * 8;
            hip_check(hipMalloc((void**)&l->d_stats, l->bricks * sizeof(float4)), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_occ16, l->bricks), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_feat, (size_t)l->rows * kLFeatures * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_y, (size_t)l->rows * 2 * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_table, l->rows), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_err, (size_t)l->rows * 4 * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_visits, (size_t)l->rows * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_alpha, 4096 * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_alpha_prefix, 4097 * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_nz_prefix, 4097 * 4), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_alpha_m, (size_t)kLMaxMult * 256 * sizeof(float)), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_sse, 64), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_flag, 64), "hipMalloc");
            hip_check(hipMalloc((void**)&l->d_count, 64), "hipMalloc");
            hip_check(hipStreamCreateWithFlags(&l->side, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&l->fork, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&l->join, hipEventDisableTiming), "hipEventCreate");
            hipStream_t st = ctx_stream(ctx);
            launch_brick_stats(vol->d_linear, vol->desc.dtype, vol->desc.dims, l->nb, l->d_stats, st);
            launch_fill_table(l->d_table, l->rows, 0, st);
            hip_check(hipGetLastError(), "brick stats launch");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            free_learned(l);
            throw;
        }
        *out = l;
    });
}

int hipdvr_learned_destroy(hipdvr_learned* l) {
    if (!l) return HIPDVR_OK;
    (void)hipStreamSynchronize(ctx_stream(l->ctx));
    free_learned(l);
    return HIPDVR_OK;
}

int hipdvr_learned_prepare(hipdvr_learned* l, const hipdvr_tf* tf, const hipdvr_params* prm, hipdvr_mlp* mlp,
                           const hipdvr_learned_opts* o) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        if (!tf || !prm) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (tf->ctx != l->ctx || (mlp && mlp->ctx != l->ctx)) fail(HIPDVR_ERR_INVALID, "handle from another context");
        validate_params(prm, tf);
        if (prm->classification != HIPDVR_CLASS_TF)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping needs the linear filter and TF classification");
        const hipdvr_learned_opts opts = default_opts(o);
        ctx_set_device(l->ctx);
        hipStream_t st = ctx_stream(l->ctx);
        upload_tf_tables(l, tf, prm->step, st);
        FeatureArgs fa;
        fa.stats = l->d_stats;
        fa.rows = l->rows;
        fa.rwm_scale = l->vol->desc.rwm_scale;
        fa.rwm_offset = l->vol->desc.rwm_offset;
        fa.dom_lo = tf->domain_lo;
        fa.dom_hi = tf->domain_hi;
        fa.tf_width = (int32_t)tf->width;
        fa.alpha = l->d_alpha;
        fa.alpha_prefix = l->d_alpha_prefix;
        fa.nz_prefix = l->d_nz_prefix;
        fa.x = l->d_feat;
        fa.occ16 = l->d_occ16;
        launch_features(fa, st);
        if (mlp) {
            launch_mlp_forward(mlp->dev, l->d_feat, l->rows, l->d_y, st);
            launch_table(l->d_y, l->rows, l->d_table, opts.skip_logit, opts.max_log2_step, st);
        } else {
            launch_fill_table(l->d_table, l->rows, 0, st);
        }
        launch_merge_occ(l->d_table, l->d_occ16, l->rows, st);
        launch_table_gain(l->d_table, l->d_occ16, l->rows, l->d_count, st);
        unsigned int gain_rows = 0;
        hip_check(hipMemcpyAsync(&gain_rows, l->d_count, 4, hipMemcpyDeviceToHost, st), "download");
        hip_check(hipGetLastError(), "learned prepare launch");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        l->prepared_tf = tf;
        l->gains = gain_rows != 0;
    });
}

int hipdvr_learned_trace_reset(hipdvr_learned* l) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        ctx_set_device(l->ctx);
        hipStream_t st = ctx_stream(l->ctx);
        hip_check(hipMemsetAsync(l->d_err, 0, (size_t)l->rows * 16, st), "memset");
        hip_check(hipMemsetAsync(l->d_visits, 0, (size_t)l->rows * 4, st), "memset");
    });
}

int hipdvr_learned_trace(hipdvr_learned* l, const hipdvr_tf* tf, const hipdvr_params* prm, const float* d_entry,
                         const float* d_exit, uint32_t w, uint32_t h) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        if (!tf || !prm || !d_entry || !d_exit || !w || !h) fail(HIPDVR_ERR_INVALID, "bad arguments");
        validate_params(prm, tf);
        if (prm->filter != HIPDVR_FILTER_LINEAR || prm->classification != HIPDVR_CLASS_TF)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping needs the linear filter and TF classification");
        // the trace reads the tiled copy through 32-bit buffer offsets
        const uint64_t vsz = l->vol->desc.dtype == HIPDVR_U8 ? 1 : 2;
        if (l->vol->tiled_elems * vsz > 0xffffff00ull)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned traces need a volume below 4 GiB; train on a smaller volume "
                                         "of the same kind (the network reads brick statistics only)");
        ctx_set_device(l->ctx);
        LearnedArgs la;
        fill_learned_args(la, l, tf, prm, d_entry, d_exit, nullptr, w, h);
        launch_trace(la, l->d_err, l->d_visits, ctx_stream(l->ctx));
        hip_check(hipGetLastError(), "trace launch");
    });
}

int hipdvr_learned_collect(hipdvr_learned* l, float tau_skip, float tau_step, float* d_x, float* d_t,
                           uint32_t capacity, uint32_t* count) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        if (!d_x || !d_t || !count) fail(HIPDVR_ERR_INVALID, "NULL argument");
        ctx_set_device(l->ctx);
        hipStream_t st = ctx_stream(l->ctx);
        const unsigned int start = *count;
        hip_check(hipMemcpyAsync(l->d_count, &start, 4, hipMemcpyHostToDevice, st), "upload");
        launch_targets(l->d_err, l->d_visits, l->d_feat, l->rows, tau_skip, tau_step, d_x, d_t, l->d_count,
                       capacity, st);
        hip_check(hipGetLastError(), "targets launch");
        unsigned int n = 0;
        hip_check(hipMemcpyAsync(&n, l->d_count, 4, hipMemcpyDeviceToHost, st), "download");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *count = n < capacity ? n : capacity;
    });
}

int hipdvr_render_learned_device(hipdvr_learned* l, const hipdvr_tf* tf, const float* d_entry,
                                 const float* d_exit, uint32_t w, uint32_t h, const hipdvr_params* prm,
                                 const hipdvr_learned_opts* o, float* d_out, hipdvr_learned_stats* stats) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        if (!tf || !prm || !d_entry || !d_exit || !d_out || !w || !h) fail(HIPDVR_ERR_INVALID, "bad arguments");
        validate_params(prm, tf);
        if (prm->filter != HIPDVR_FILTER_LINEAR || prm->classification != HIPDVR_CLASS_TF)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping needs the linear filter and TF classification");
        if (prm->shading != HIPDVR_SHADE_NONE && prm->shading != HIPDVR_SHADE_PHONG)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping shades with Phong or not at all");
        if (prm->shading != HIPDVR_SHADE_NONE && prm->gradient != HIPDVR_GRADIENT_CENTRAL)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping shades with central-difference gradients");
        if (prm->compositing != HIPDVR_COMPOSITE_DVR)
            fail(HIPDVR_ERR_UNSUPPORTED, "learned stepping composites DVR");
        const hipdvr_learned_opts opts = default_opts(o);
        ctx_set_device(l->ctx);
        hipStream_t st = ctx_stream(l->ctx);
        hipdvr_volume* vol = const_cast<hipdvr_volume*>(l->vol);
        ensure_occupancy(l->ctx, vol, tf);
        const SparseGrid sg = sparse_grid(l, w, h, (uint32_t)opts.remarch_stride);
        const uint32_t stride = sg.stride, offs = sg.offs;
        const int swc = sg.w, shc = sg.h;
        const bool phong = prm->shading == HIPDVR_SHADE_PHONG;
        const int shade = phong ? (prm->exact_shading ? 2 : 1) : 0;
        const bool count = stats != nullptr;
        if (count) launch_zero_u64(l->ctx->d_counters, kNumCounters, st);
        if (!l->gains) {
            // the table keeps every occupied brick at the base step: the
            // learned frame could only approximate the plain frame, at a
            // higher cost, so the plain frame is rendered (no judge needed)
            const bool timed0 = ctx_timing_start(l->ctx);
            ensure_cells(vol, st);
            MarchArgs pa;
            fill_march_args(pa, vol, tf, prm, d_entry, d_exit, d_out, w, h, count ? l->ctx->d_counters : nullptr);
            launch_march_cell(pa, vol->desc.dtype, shade, count, vol->cells_wide, default_batch(shade), st);
            hip_check(hipGetLastError(), "learned render launch");
            if (timed0) ctx_timing_stop(l->ctx);
            if (count) {
                unsigned long long c[kNumCounters];
                hip_check(hipMemcpyAsync(c, l->ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st), "download");
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
                stats->samples = c[2];   // reconstructed samples, as the learned march counts them
                stats->shaded = c[3];
                stats->sparse_psnr = 999.f;
                stats->fell_back = 0;
            }
            return;
        }
        const bool timed = ctx_timing_start(l->ctx);
        // 1. sparse plain re-march (the plain march of the judge's pixels,
        //    sparse_grid) on the side stream, beside 2. the learned frame
        ensure_cells(vol, st);
        hip_check(hipEventRecord(l->fork, st), "hipEventRecord");
        hip_check(hipStreamWaitEvent(l->side, l->fork, 0), "hipStreamWaitEvent");
        launch_sparse(l, vol, tf, prm, d_entry, d_exit, w, h, sg, shade, (float*)l->d_sparse, l->side);
        hip_check(hipEventRecord(l->join, l->side), "hipEventRecord");
        LearnedArgs la;
        fill_learned_args(la, l, tf, prm, d_entry, d_exit, d_out, w, h);
        la.m.cells = vol->d_cells;
        launch_learned(la, phong, count, vol->cells_wide, st);
        hip_check(hipStreamWaitEvent(st, l->join, 0), "hipStreamWaitEvent");
        // 3. judge: flag = PSNR below the floor
        const float mse_max = powf(10.f, -opts.psnr_floor / 10.f);
        launch_judge((const float4*)d_out, (int)w, (int)h, l->d_sparse, swc, shc, (int)stride, (int)sg.block,
                     (int)offs,
                     l->d_sse, l->d_flag, mse_max, st);
        // 4. fallback: the plain frame, launched always, runs only if flagged
        MarchArgs fb;
        fill_march_args(fb, vol, tf, prm, d_entry, d_exit, d_out, w, h, nullptr);
        fb.cond_flag = l->d_flag;
        launch_march_cell(fb, vol->desc.dtype, shade, false, vol->cells_wide, default_batch(shade), st);
        hip_check(hipGetLastError(), "learned render launch");
        if (timed) ctx_timing_stop(l->ctx);
        if (count) {
            unsigned long long c[kNumCounters];
            float sse = 0.f;
            int flag = 0;
            hip_check(hipMemcpyAsync(c, l->ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st), "download");
            hip_check(hipMemcpyAsync(&sse, l->d_sse, 4, hipMemcpyDeviceToHost, st), "download");
            hip_check(hipMemcpyAsync(&flag, l->d_flag, 4, hipMemcpyDeviceToHost, st), "download");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            const double mse = (double)sse / ((double)swc * shc * 4);
            stats->samples = c[1];
            stats->shaded = c[3];
            stats->sparse_psnr = mse > 0 ? (float)(10.0 * std::log10(1.0 / mse)) : 999.f;
            stats->fell_back = flag;
        }
    });
}

int hipdvr_learned_sparse_dims(uint32_t w, uint32_t h, const hipdvr_learned_opts* o, uint32_t* sw, uint32_t* sh) {
    if (!w || !h || !sw || !sh) return HIPDVR_ERR_INVALID;
    const hipdvr_learned_opts opts = default_opts(o);
    const SparseGrid g = sparse_dims(w, h, (uint32_t)opts.remarch_stride);
    *sw = (uint32_t)g.w;
    *sh = (uint32_t)g.h;
    return HIPDVR_OK;
}

int hipdvr_learned_sparse_remarch(hipdvr_learned* l, const hipdvr_tf* tf, const float* d_entry,
                                  const float* d_exit, uint32_t w, uint32_t h, const hipdvr_params* prm,
                                  const hipdvr_learned_opts* o, float* d_sparse, uint64_t capacity, uint32_t* sw,
                                  uint32_t* sh) {
    if (!l) return HIPDVR_ERR_INVALID;
    return ctx_guarded(l->ctx, [&] {
        if (!tf || !prm || !d_entry || !d_exit || !d_sparse || !sw || !sh || !w || !h)
            fail(HIPDVR_ERR_INVALID, "bad arguments");
        validate_params(prm, tf);
        if (prm->filter != HIPDVR_FILTER_LINEAR || prm->classification != HIPDVR_CLASS_TF ||
            prm->compositing != HIPDVR_COMPOSITE_DVR ||
            (prm->shading != HIPDVR_SHADE_NONE && prm->shading != HIPDVR_SHADE_PHONG) ||
            (prm->shading != HIPDVR_SHADE_NONE && prm->gradient != HIPDVR_GRADIENT_CENTRAL))
            fail(HIPDVR_ERR_UNSUPPORTED, "the sparse re-march serves the learned path's options");
        const hipdvr_learned_opts opts = default_opts(o);
        const SparseGrid need = sparse_dims(w, h, (uint32_t)opts.remarch_stride);
        if ((uint64_t)need.w * (uint64_t)need.h > capacity)
            fail(HIPDVR_ERR_INVALID, "d_sparse holds " + std::to_string(capacity) + " pixels, the grid needs " +
                                         std::to_string(need.w) + " x " + std::to_string(need.h));
        ctx_set_device(l->ctx);
        hipStream_t st = ctx_stream(l->ctx);
        hipdvr_volume* vol = const_cast<hipdvr_volume*>(l->vol);
        ensure_occupancy(l->ctx, vol, tf);
        ensure_cells(vol, st);
        const SparseGrid g = sparse_grid(l, w, h, (uint32_t)opts.remarch_stride);
        const int shade = prm->shading == HIPDVR_SHADE_PHONG ? (prm->exact_shading ? 2 : 1) : 0;
        launch_sparse(l, vol, tf, prm, d_entry, d_exit, w, h, g, shade, d_sparse, st);
        hip_check(hipGetLastError(), "sparse re-march launch");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *sw = (uint32_t)g.w;
        *sh = (uint32_t)g.h;
    });
}

}  // extern "C"
