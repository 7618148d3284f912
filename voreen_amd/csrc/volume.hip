// volume.hip -- volume upload: linear copy, bricked copy, brick min/max.
#include <hip/hip_runtime.h>

#include <string>

#include "errors.h"
#include "volume.h"

namespace hipdvr {

hipdvr_volume* volume_create(hipdvr_ctx* ctx, hipStream_t st, const hipdvr_volume_desc* d,
                             const void* voxels, bool on_device) {
    for (int i = 0; i < 3; ++i)
        if (d->dims[i] < 1 || d->dims[i] > (1u << 20))
            fail(HIPDVR_ERR_INVALID, "volume dimensions must be in [1, 2^20]");
    if (d->dtype != HIPDVR_U8 && d->dtype != HIPDVR_U16)
        fail(HIPDVR_ERR_INVALID, "dtype must be HIPDVR_U8 or HIPDVR_U16");
    uint32_t bl = d->brick_log2 ? d->brick_log2 : 4;
    if (bl < 3 || bl > 5) fail(HIPDVR_ERR_INVALID, "brick_log2 must be 3, 4 or 5");
    hipdvr_volume* v = new hipdvr_volume();
    v->ctx = ctx;
    for (int i = 0; i < 3; ++i) {
        v->desc.dims[i] = d->dims[i];
        v->desc.spacing[i] = d->spacing[i];
        v->desc.offset[i] = d->offset[i];
        v->nbricks[i] = (d->dims[i] + (1u << bl) - 1) >> bl;
    }
    v->desc.dtype = d->dtype;
    v->desc.rwm_scale = d->rwm_scale;
    v->desc.rwm_offset = d->rwm_offset;
    v->desc.brick_log2 = bl;
    const size_t vsz = d->dtype == HIPDVR_U8 ? 1 : 2;
    const size_t bytes = (size_t)d->dims[0] * d->dims[1] * d->dims[2] * vsz;
    try {
        hip_check(hipMalloc(&v->d_linear, bytes), "hipMalloc(volume)");
        hip_check(hipMemcpyAsync(v->d_linear, voxels, bytes,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
               "volume upload");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    } catch (...) {
        volume_destroy(v);
        throw;
    }
    return v;
}

void volume_destroy(hipdvr_volume* v) {
    if (!v) return;
    if (v->d_linear) (void)hipFree(v->d_linear);
    if (v->d_bricks) (void)hipFree(v->d_bricks);
    if (v->d_minmax) (void)hipFree(v->d_minmax);
    delete v;
}

}  // namespace hipdvr
