// learned_train.cpp -- hipdvr_learned_train: the stepping network's whole
// training loop behind the C-ABI, so that a C++ host (the HIPRaycaster
// processor) can train without Python.
//
// For each of n_tfs seeded random transfer functions (4-8 keys) and the
// caller's extra tables, the plain marcher traces the learned state's volume
// from n_views orbit cameras at trace_size^2 (hipdvr_learned_trace); the
// visited (brick, octant) rows become training rows (hipdvr_learned_collect),
// and the network takes Adam steps on random contiguous minibatches of them
// (hipdvr_mlp_train_step).  Every step runs the library's own kernels; this
// file only sequences them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hipdvr.h"
#include "../../include/hipdvr_learned.h"
#include "ctx.h"
#include "errors.h"
#include "objects.h"
#include "volume.h"

struct hipdvr_learned;

namespace hipdvr {

hipdvr_ctx* learned_ctx(const hipdvr_learned* l);
const hipdvr_volume* learned_volume(const hipdvr_learned* l);

namespace {

// SplitMix64 (the generator of the seeded workloads).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t randint(uint32_t n) { return (uint32_t)(next() % n); }
};

// 4-8 keys: intensities U[0,1] sorted, RGBA8 U{0..255}.
std::vector<hipdvr_tf_key> random_tf_keys(uint64_t seed) {
    SplitMix64 r(seed);
    const uint32_t n = 4 + r.randint(5);
    SplitMix64 k(r.next());
    std::vector<float> its(n);
    for (auto& v : its) v = (float)k.uniform();
    std::sort(its.begin(), its.end());
    std::vector<hipdvr_tf_key> keys(n);
    for (uint32_t i = 0; i < n; ++i) {
        keys[i].intensity = its[i];
        keys[i].split = 0;
        for (int c = 0; c < 4; ++c) keys[i].color_l[c] = keys[i].color_r[c] = (uint8_t)k.randint(256);
    }
    return keys;
}

// Orbit camera of view v of n around the origin (radius 3.5, fovy 30), at an
// elevation cycling over 10, 26.7, 43.3 and 60 degrees.
hipdvr_camera orbit_camera(uint32_t v, uint32_t n) {
    const double th = 2.0 * 3.14159265358979323846 * v / n;
    const double el = (10.0 + 50.0 * (double)(v % 4) / 3.0) * 3.14159265358979323846 / 180.0;
    hipdvr_camera c;
    c.eye[0] = (float)(3.5 * std::sin(th) * std::cos(el));
    c.eye[1] = (float)(3.5 * std::sin(el));
    c.eye[2] = (float)(3.5 * std::cos(th) * std::cos(el));
    c.focus[0] = c.focus[1] = c.focus[2] = 0.f;
    c.up[0] = 0.f;
    c.up[1] = 1.f;
    c.up[2] = 0.f;
    c.fovy_deg = 30.f;
    c.near_dist = 0.1f;
    return c;
}

hipdvr_learned_train_opts default_train_opts(const hipdvr_learned_train_opts* o) {
    hipdvr_learned_train_opts d;
    std::memset(&d, 0, sizeof d);
    d.n_tfs = 200;
    d.n_views = 16;
    d.trace_size = 192;
    d.sampling_rate = 2.f;
    d.tau_skip = 3e-3f;
    d.tau_step = 8e-3f;
    d.min_steps = 2000;
    d.epochs = 4;
    d.batch = 8192;
    d.lr = 2e-3f;
    d.seed = 0x5EED;
    if (!o) return d;
    if (o->n_tfs) d.n_tfs = o->n_tfs;
    if (o->n_views) d.n_views = o->n_views;
    if (o->trace_size) d.trace_size = o->trace_size;
    if (o->sampling_rate > 0.f) d.sampling_rate = o->sampling_rate;
    if (o->tau_skip > 0.f) d.tau_skip = o->tau_skip;
    if (o->tau_step > 0.f) d.tau_step = o->tau_step;
    if (o->min_steps) d.min_steps = o->min_steps;
    if (o->epochs) d.epochs = o->epochs;
    if (o->batch) d.batch = o->batch;
    if (o->lr > 0.f) d.lr = o->lr;
    if (o->seed) d.seed = o->seed;
    d.extra_tfs = o->extra_tfs;
    d.n_extra = o->n_extra;
    d.extra_width = o->extra_width;
    d.max_rows = o->max_rows;
    return d;
}

// Device buffer freed on scope exit.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

void rc_check(hipdvr_ctx* ctx, int rc, const char* what) {
    if (rc != HIPDVR_OK) fail(rc, std::string(what) + ": " + hipdvr_last_error(ctx));
}

}  // namespace
}  // namespace hipdvr

using namespace hipdvr;

extern "C" int hipdvr_learned_train(hipdvr_learned* l, hipdvr_mlp* mlp, const hipdvr_learned_train_opts* opts,
                                    hipdvr_learned_train_info* info) {
    if (!l || !mlp) return HIPDVR_ERR_INVALID;
    hipdvr_ctx* ctx = learned_ctx(l);
    const hipdvr_volume* vol = learned_volume(l);
    // the nested C-ABI calls below record their own errors in ctx; the
    // outer guard keeps the first failure's message
    return ctx_guarded(ctx, [&] {
        const auto t0 = std::chrono::steady_clock::now();
        const hipdvr_learned_train_opts o = default_train_opts(opts);
        if (o.trace_size > 4096 || o.n_views > 4096 || o.n_tfs > 100000 || o.batch > (1u << 22))
            fail(HIPDVR_ERR_INVALID, "training options out of range");
        if (o.n_extra && !o.extra_tfs) fail(HIPDVR_ERR_INVALID, "extra_tfs is NULL");
        const VolumeDesc& desc = vol->desc;
        const uint32_t tf_width = desc.dtype == HIPDVR_U8 ? 256u : 4096u;
        const uint32_t extra_width = o.extra_width ? o.extra_width : tf_width;
        if (extra_width < 2 || extra_width > 4096) fail(HIPDVR_ERR_INVALID, "extra_width must be in [2, 4096]");
        const uint32_t* dims = desc.dims;
        const uint32_t md = std::max(dims[0], std::max(dims[1], dims[2]));
        const float step = 1.f / ((float)md * o.sampling_rate);
        const uint64_t rows_per_tf =
            (uint64_t)((dims[0] + 15) / 16) * ((dims[1] + 15) / 16) * ((dims[2] + 15) / 16) * 8ull;
        const uint64_t cap_limit = o.max_rows ? o.max_rows : 24000000ull;
        const uint32_t capacity = (uint32_t)std::min<uint64_t>((uint64_t)(o.n_tfs + o.n_extra) * rows_per_tf, cap_limit);
        if (capacity == 0) fail(HIPDVR_ERR_INVALID, "nothing to train on");

        DevBuf dx, dt, de, dxx;
        hip_check(hipMalloc(&dx.p, (size_t)capacity * HIPDVR_MLP_INPUTS * sizeof(float)), "hipMalloc(training x)");
        hip_check(hipMalloc(&dt.p, (size_t)capacity * HIPDVR_MLP_TARGETS * sizeof(float)), "hipMalloc(training t)");
        const size_t img = (size_t)o.trace_size * o.trace_size * 4 * sizeof(float);
        hip_check(hipMalloc(&de.p, img * o.n_views), "hipMalloc(trace entry)");
        hip_check(hipMalloc(&dxx.p, img * o.n_views), "hipMalloc(trace exit)");
        // entry/exit of every trace view: the proxy box centred at the origin,
        // longest extent 1 (unit spacing)
        float llf[3], urb[3];
        for (int i = 0; i < 3; ++i) {
            const float ext = (float)dims[i] / (float)md;
            llf[i] = -ext / 2;
            urb[i] = ext / 2;
        }
        for (uint32_t v = 0; v < o.n_views; ++v) {
            const hipdvr_camera cam = orbit_camera(v, o.n_views);
            rc_check(ctx, hipdvr_entry_exit_device(ctx, &cam, llf, urb, o.trace_size, o.trace_size, 0,
                                                   (float*)((char*)de.p + img * v), (float*)((char*)dxx.p + img * v)),
                     "hipdvr_entry_exit_device");
        }
        hipdvr_params prm;
        std::memset(&prm, 0, sizeof prm);
        prm.step = step;
        prm.filter = HIPDVR_FILTER_LINEAR;
        prm.classification = HIPDVR_CLASS_TF;
        prm.compositing = HIPDVR_COMPOSITE_DVR;
        prm.shading = HIPDVR_SHADE_NONE;
        prm.shininess = 60.f;
        prm.kernel = HIPDVR_KERNEL_AUTO;
        uint32_t count = 0;
        std::vector<uint8_t> table(4 * (size_t)tf_width);
        const uint32_t n_tables = o.n_tfs + o.n_extra;
        uint32_t traced = 0;
        for (uint32_t k = 0; k < n_tables && count < capacity; ++k) {
            const uint8_t* rgba = nullptr;
            uint32_t width = tf_width;
            if (k < o.n_tfs) {
                const std::vector<hipdvr_tf_key> keys = random_tf_keys(o.seed + 7919ull * k);
                rc_check(ctx, hipdvr_tf_from_keys(keys.data(), (int32_t)keys.size(), (int32_t)tf_width, 1.f, 0.f, 1.f,
                                                  HIPDVR_ALPHA_USE, table.data()),
                         "hipdvr_tf_from_keys");
                rgba = table.data();
            } else {
                width = extra_width;
                rgba = o.extra_tfs + (size_t)(k - o.n_tfs) * 4 * extra_width;
            }
            hipdvr_tf* tf = nullptr;
            rc_check(ctx, hipdvr_tf_create(ctx, rgba, width, 0.f, 1.f, &tf), "hipdvr_tf_create");
            int rc = hipdvr_learned_prepare(l, tf, &prm, nullptr, nullptr);
            if (rc == HIPDVR_OK) rc = hipdvr_learned_trace_reset(l);
            for (uint32_t v = 0; v < o.n_views && rc == HIPDVR_OK; ++v)
                rc = hipdvr_learned_trace(l, tf, &prm, (const float*)((char*)de.p + img * v),
                                          (const float*)((char*)dxx.p + img * v), o.trace_size, o.trace_size);
            if (rc == HIPDVR_OK)
                rc = hipdvr_learned_collect(l, o.tau_skip, o.tau_step, (float*)dx.p, (float*)dt.p, capacity, &count);
            const std::string err = rc == HIPDVR_OK ? std::string() : std::string(hipdvr_last_error(ctx));
            hipdvr_tf_destroy(tf);
            if (rc != HIPDVR_OK) fail(rc, "training trace: " + err);
            ++traced;
        }
        if (count == 0) fail(HIPDVR_ERR_INVALID, "the traces visited no bricks");
        const auto t1 = std::chrono::steady_clock::now();
        const uint32_t rows = std::min(o.batch, count);
        const uint64_t want = (uint64_t)o.epochs * count / o.batch;
        const uint32_t steps = (uint32_t)std::max<uint64_t>(o.min_steps, want);
        SplitMix64 r(o.seed ^ 0xA5A5A5A5ull);
        const uint32_t span = count > o.batch ? count - o.batch : 1;
        float first = 0.f, last = 0.f;
        for (uint32_t i = 0; i < steps; ++i) {
            const uint32_t off = r.randint(span);
            const bool want_loss = i == 0 || i + 1 == steps;
            float loss = 0.f;
            rc_check(ctx,
                     hipdvr_mlp_train_step(mlp, (const float*)dx.p + (size_t)off * HIPDVR_MLP_INPUTS,
                                           (const float*)dt.p + (size_t)off * HIPDVR_MLP_TARGETS, rows, o.lr,
                                           want_loss ? &loss : nullptr),
                     "hipdvr_mlp_train_step");
            if (i == 0) first = loss;
            if (i + 1 == steps) last = loss;
        }
        rc_check(ctx, hipdvr_synchronize(ctx), "hipdvr_synchronize");
        const auto t2 = std::chrono::steady_clock::now();
        if (info) {
            info->rows = count;
            info->tfs = traced;
            info->views = o.n_views;
            info->trace_size = o.trace_size;
            info->steps = steps;
            info->loss_first = first;
            info->loss_last = last;
            info->trace_seconds = std::chrono::duration<float>(t1 - t0).count();
            info->seconds = std::chrono::duration<float>(t2 - t0).count();
        }
    });
}
