// volume.h -- device representation of a volume (internal).
//
// HBM layout:
//   d_linear  x-fastest copy, z*X*Y + y*X + x (the VolumeRAM layout,
//             include/voreen/core/datastructures/volume/volumeatomic.h:599-608)
//   d_bricks  B^3 bricks (B = 2^brick_log2) stored with a +1 apron on the
//             upper side of every axis, (B+1)^3 voxels per brick, bricks in
//             Morton (Z-order) of their brick coordinates.  Every trilinear
//             footprint whose lower corner lies in a brick is inside that
//             brick's apron-extended copy.
//   d_minmax  per brick (min, max) over the apron-extended voxels, as raw
//             voxel values (uint32 pairs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipdvr.h"

namespace hipdvr {

struct VolumeDesc {
    uint32_t dims[3];
    int32_t dtype;
    float rwm_scale, rwm_offset;
    float spacing[3], offset[3];
    uint32_t brick_log2;
};

}  // namespace hipdvr

struct hipdvr_volume {
    hipdvr_ctx* ctx = nullptr;
    hipdvr::VolumeDesc desc{};
    void* d_linear = nullptr;
    void* d_bricks = nullptr;
    uint2* d_minmax = nullptr;
    uint32_t nbricks[3] = {0, 0, 0};
    uint64_t brick_count = 0;
    uint64_t brick_stride = 0;   // voxels per stored brick ((B+1)^3, padded)
};

namespace hipdvr {

hipdvr_volume* volume_create(hipdvr_ctx* ctx, hipStream_t st, const hipdvr_volume_desc* desc,
                             const void* voxels, bool on_device);
void volume_destroy(hipdvr_volume* vol);

}  // namespace hipdvr
