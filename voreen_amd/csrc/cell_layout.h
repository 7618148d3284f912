// cell_layout.h -- where the cell record of cell (x, y, z) lives.
//
// A record holds the 8 footprint voxels of one cell (volume.hip builds them,
// march_cell.h reads them).  Records come in groups of 2x2x2 cells (one
// 128-byte line for uint16), the parities of x, y, z in bits 0, 1, 2 of the
// record index.  Groups come in bricks of B^3 groups (B = 2^kCellBrickLog2,
// 2B voxels per edge), groups x-fastest inside a brick, bricks x-fastest;
// each axis is padded to whole bricks.  The index is a sum of one term per
// axis:
//   term_i(l) = ((l >> 1) >> log2 B) * brick_stride_i
//             + (((l >> 1) & (B - 1)) << (3 + i * log2 B)) + ((l & 1) << i)
// with brick strides B^3*8, NBx*B^3*8 and NBx*NBy*B^3*8 (NB = bricks per axis).
// A ray then stays inside one brick's contiguous 2B^3*8 records (512 KiB of
// uint16 records at B = 16) for 2B voxels of travel along any axis, instead of
// moving to a new group slab -- X*Y/4 records away -- every two voxels of z.
// kCellBrickLog2 = 0 is the plain x-fastest order of groups.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hipdvr {

#ifndef HIPDVR_CELL_BRICK_LOG2
#define HIPDVR_CELL_BRICK_LOG2 4
#endif
#ifndef HIPDVR_CELL_MORTON
#define HIPDVR_CELL_MORTON 0
#endif
constexpr uint32_t kCellBrickLog2 = HIPDVR_CELL_BRICK_LOG2;
// groups inside a brick in Morton (z-order) instead of x-fastest order
constexpr bool kCellMorton = HIPDVR_CELL_MORTON != 0;
constexpr uint32_t kCellBrick = 1u << kCellBrickLog2;          // groups per brick edge
constexpr uint32_t kCellBrickMask = kCellBrick - 1;
constexpr uint32_t kCellBrickRecLog2 = 3 * kCellBrickLog2 + 3;  // records per brick, log2

__host__ __device__ inline uint32_t cell_groups(uint32_t dim) { return (dim + 1) >> 1; }
__host__ __device__ inline uint32_t cell_bricks(uint32_t dim) {
    return (cell_groups(dim) + kCellBrickMask) >> kCellBrickLog2;
}
// records of a volume (padding included)
__host__ __device__ inline uint64_t cell_record_count(const uint32_t dims[3]) {
    return ((uint64_t)cell_bricks(dims[0]) * cell_bricks(dims[1]) * cell_bricks(dims[2])) << kCellBrickRecLog2;
}
// brick stride of axis i, in records
__host__ __device__ inline uint64_t cell_brick_stride(const uint32_t dims[3], int i) {
    uint64_t s = (uint64_t)1 << kCellBrickRecLog2;
    if (i > 0) s *= cell_bricks(dims[0]);
    if (i > 1) s *= cell_bricks(dims[1]);
    return s;
}
// in-brick stride of axis i between neighbouring groups, in records
__host__ __device__ constexpr uint32_t cell_group_stride(int i) { return 1u << (3 + i * kCellBrickLog2); }

// Morton order inside a brick: bit k of the group's place along axis i goes
// to bit 3k + i of the group's in-brick number.
__host__ __device__ __forceinline__ uint32_t cell_spread(uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCellBrickLog2; ++k) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}
__host__ __device__ __forceinline__ uint32_t cell_compact(uint32_t m) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCellBrickLog2; ++k) r |= ((m >> (3 * k)) & 1u) << k;
    return r;
}
// bits of axis i in a record index (group number shifted past the parities)
__host__ __device__ constexpr uint32_t cell_axis_mask(int i) {
    uint32_t m = 0;
    for (uint32_t k = 0; k < kCellBrickLog2; ++k) m |= 1u << (3 * k + i);
    return m << 3;
}

// Record index of cell lo (nbx, nby: bricks along x and y).
template <typename off_t>
__host__ __device__ __forceinline__ off_t cell_index(const uint32_t lo[3], uint32_t nbx, uint32_t nby) {
    uint32_t g[3], local = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        g[i] = lo[i] >> 1;
        const uint32_t o = g[i] & kCellBrickMask;
        local |= (kCellMorton ? cell_spread(o) << (3 + i) : o << (3 + i * kCellBrickLog2)) | ((lo[i] & 1u) << i);
    }
    const uint32_t brick = ((g[2] >> kCellBrickLog2) * nby + (g[1] >> kCellBrickLog2)) * nbx +
                           (g[0] >> kCellBrickLog2);
    return ((off_t)brick << kCellBrickRecLog2) + (off_t)local;
}

// Records of the cells lo - e_i and lo + e_i from the record index `own` of
// cell lo alone (its bits hold lo's parity and its group's place in the
// brick along i), brick_stride = cell_brick_stride(i); both cells must lie
// inside the volume.
template <typename off_t>
__host__ __device__ __forceinline__ void cell_neighbours(off_t own, int i, off_t brick_stride, off_t& lower,
                                                         off_t& upper) {
    const off_t bit = (off_t)1 << i;
    if (kCellMorton) {
        // step the axis bits of the group number up or down (carries and
        // borrows run through the other axes' bits), or wrap them to the
        // facing group of the neighbouring brick
        const uint32_t M = cell_axis_mask(i), o = (uint32_t)own, ax = o & M;
        if (o & (1u << i)) {
            lower = own - bit;
            const off_t base = own - bit - (off_t)ax;
            upper = ax == M ? base + brick_stride : base + (off_t)(((o | ~M) + 8u) & M);
        } else {
            upper = own + bit;
            const off_t base = own + bit - (off_t)ax;
            lower = ax == 0 ? base + (off_t)M - brick_stride : base + (off_t)((ax - 8u) & M);
        }
        return;
    }
    const uint32_t gl = ((uint32_t)own >> (3 + i * kCellBrickLog2)) & kCellBrickMask;
    const off_t gs = (off_t)cell_group_stride(i);                    // next group, same brick
    const off_t cross = brick_stride - (off_t)kCellBrickMask * gs;   // last group to the next brick's first
    if ((uint32_t)own & (1u << i)) {
        lower = own - bit;                                        // same group
        upper = own - bit + (gl != kCellBrickMask ? gs : cross);  // next group's even cell
    } else {
        lower = own + bit - (gl != 0 ? gs : cross);               // previous group's odd cell
        upper = own + bit;                                        // same group
    }
}

// Cell (x, y, z) of record i (the inverse of the index; cells of the padding
// come out beyond the volume's dims).
__host__ __device__ inline void cell_of_record(uint64_t i, const uint32_t dims[3], uint32_t c[3]) {
    const uint32_t par = (uint32_t)(i & 7u);
    const uint32_t local = (uint32_t)((i >> 3) & (((uint64_t)1 << (3 * kCellBrickLog2)) - 1));
    const uint32_t brick = (uint32_t)(i >> kCellBrickRecLog2);   // < 2^32 bricks (32-bit arithmetic)
    const uint32_t nbx = cell_bricks(dims[0]), nby = cell_bricks(dims[1]);
    const uint32_t bxy = brick / nbx;
    const uint32_t b[3] = {brick - bxy * nbx, bxy % nby, bxy / nby};
    for (int k = 0; k < 3; ++k) {
        const uint32_t o = kCellMorton ? cell_compact(local >> k) : (local >> (k * kCellBrickLog2)) & kCellBrickMask;
        c[k] = ((b[k] << kCellBrickLog2) + o) * 2u + ((par >> k) & 1u);
    }
}

}  // namespace hipdvr
