// volume.h -- device representation of a volume (internal).
//
// HBM layout of one volume:
//   d_linear  x-fastest copy, z*X*Y + y*X + x (the VolumeRAM layout,
//             include/voreen/core/datastructures/volume/volumeatomic.h:599-608);
//             read by the reference-layout kernel.
//   d_tiled   the same voxels in 4x4x4 blocks: dims padded up to multiples of 4,
//             blocks in x-fastest block order, voxels x-fastest inside a block.
//             One block of uint16 is 128 bytes = one cache line, so the 8
//             corners of a trilinear footprint touch 1..8 lines whatever the
//             ray direction.  The element offset is separable per axis:
//               off(x,y,z) = ox(x) + oy(y) + oz(z)
//               ox(x) = (x>>2)*64 + (x&3)
//               oy(y) = (y>>2)*64*NBX + (y&3)*4
//               oz(z) = (z>>2)*64*NBX*NBY + (z&3)*16
//             Padding voxels replicate the nearest edge voxel (never read:
//             every coordinate is clamped to dims-1 before the fetch).
//   d_cells   corner-replicated cell records, built on first use by the cell
//             kernel: the record of cell (x, y, z) packs the 8 voxels
//             (x..x+1, y..y+1, z..z+1), clamped to the volume, in trilinear
//             order (16 bytes for uint16, 8 for uint8).  Cells are grouped
//             2x2x2 per group (one 128-byte line for uint16), groups
//             x-fastest over the dims rounded up to even:
//               rec(x,y,z) = ((x>>1)*8 + (x&1)) + ((y>>1)*8*GX + (y&1)*2)
//                            + ((z>>1)*8*GX*GY + (z&1)*4)
//             8x the voxel bytes; bounded by kMaxCellBytes.
//   d_minmax  per brick of B^3 voxels (B = 2^brick_log2), the (min, max) raw
//             voxel value over the brick extended by one voxel on the upper
//             side of every axis (clamped to the volume): every trilinear or
//             nearest footprint whose lower corner lies in the brick is inside.
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
    void* d_tiled = nullptr;
    uint2* d_minmax = nullptr;
    uint8_t* d_occ = nullptr;        // per-brick occupancy for the last TF mapping
    uint32_t* d_occ_bits = nullptr;  // the same, one bit per brick
    uint32_t occ_words = 0;
    uint64_t occ_key[4] = {0, 0, 0, 0};  // TF identity/domain/rwm the occupancy belongs to
    uint32_t nblocks[3] = {0, 0, 0}; // 4^3 blocks per axis
    uint32_t nbricks[3] = {0, 0, 0};
    uint64_t brick_count = 0;
    uint64_t tiled_elems = 0;
    void* d_cells = nullptr;         // cell records (null until first use)
    uint64_t cell_bytes = 0;
    bool cells_wide = false;         // more than 2^32 - 256 bytes: 64-bit addressing
    void* d_grad = nullptr;          // gradient volume of the 2D TF path (same dims)
    int32_t grad_dtype = 0;          // HIPDVR_U8 / HIPDVR_U16 / HIPDVR_F32
    uint32_t grad_channels = 0;      // 3 or 4 interleaved channels per voxel
};

namespace hipdvr {

hipdvr_volume* volume_create(hipdvr_ctx* ctx, hipStream_t st, const hipdvr_volume_desc* desc,
                             const void* voxels, bool on_device);
void volume_destroy(hipdvr_volume* vol);

// Largest cell-record array the library allocates for one volume.
constexpr uint64_t kMaxCellBytes = 96ull << 30;
uint64_t cell_record_bytes(const hipdvr_volume* vol);
bool cells_allowed(const hipdvr_volume* vol);
// Brick min/max table of the volume for bricks of 2^bl voxels (nb bricks per
// axis), as built at creation for the volume's own brick size.
void launch_brick_minmax(const hipdvr_volume* vol, uint32_t bl, const uint32_t nb[3], uint2* out, hipStream_t st);
// Raw minimum and maximum voxel of the whole volume (synchronous).
void volume_raw_minmax(const hipdvr_volume* vol, hipStream_t st, uint32_t out[2]);
// Build the cell records if they are not there yet (synchronous).
void ensure_cells(hipdvr_volume* vol, hipStream_t st);

// Per-brick occupancy for a TF mapping: a brick is empty when no TF entry
// within one index of the range its [min, max] (widened by 1e-5 in
// normalized intensity) maps to has a non-zero alpha byte.
struct OccArgs {
    const uint2* minmax;
    uint8_t* occ;
    uint64_t count;
    int32_t dtype;
    float rwm_scale, rwm_offset, dom_lo, dom_hi;
    int32_t tf_width;
    const uint32_t* alpha_prefix;   // [w+1]: number of non-zero alpha bytes in [0, i)
    uint32_t* bits;                 // [ceil(count/32)] occupancy bits (written after occ)
};
void launch_occupancy(const OccArgs& a, hipStream_t st);

}  // namespace hipdvr
