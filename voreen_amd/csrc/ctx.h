// ctx.h -- the context object behind hipdvr_ctx and the guard that turns
// internal errors into C-ABI status codes (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/hipdvr.h"
#include "errors.h"

struct hipdvr_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* d_counters = nullptr;   // kNumCounters x u64
    std::string last_error;
    // per-launch timing (hipdvr_timing_begin/read)
    std::vector<hipEvent_t> tev;                 // 2 per slot
    uint32_t tcap = 0, tcount = 0;
    // staging of hipdvr_render (host entry/exit in, host image out): the
    // device copies of entry, exit and output, and pinned host buffers of
    // the same size; grown when a frame is larger, freed with the context
    float* d_stage = nullptr;
    float* h_stage = nullptr;
    size_t stage_bytes = 0;                      // each of the two, 3 images
};

namespace hipdvr {

inline hipStream_t ctx_stream(const hipdvr_ctx* c) { return c->stream; }
inline void ctx_set_device(hipdvr_ctx* c) { hip_check(hipSetDevice(c->device), "hipSetDevice"); }

template <typename F>
int ctx_guarded(hipdvr_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return HIPDVR_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed";
        return HIPDVR_ERR_NOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return HIPDVR_ERR_HIP;
    } catch (...) {
        if (ctx) ctx->last_error = "unknown error";
        return HIPDVR_ERR_HIP;
    }
}

// Record the start/end event of a timed launch (hipdvr_timing_begin/read).
inline bool ctx_timing_start(hipdvr_ctx* ctx) {
    if (ctx->tcount >= ctx->tcap) return false;
    hip_check(hipEventRecord(ctx->tev[2 * ctx->tcount], ctx->stream), "hipEventRecord");
    return true;
}
inline void ctx_timing_stop(hipdvr_ctx* ctx) {
    hip_check(hipEventRecord(ctx->tev[2 * ctx->tcount + 1], ctx->stream), "hipEventRecord");
    ++ctx->tcount;
}

}  // namespace hipdvr
