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
#include "mlp.h"

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
