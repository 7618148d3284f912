/*
 * hipdvr_synth.h -- seeded synthetic workloads for the DVR ray march.
 *
 * These generators have no counterpart in the reference (Voreen reads real
 * volumes from disk); they produce the benchmark and test volumes
 * deterministically from a seed (SplitMix64), on the host, multi-threaded,
 * or on the device (hipdvr_synth_turbulence_device).
 */
#ifndef HIPDVR_SYNTH_H
#define HIPDVR_SYNTH_H

#include <stdint.h>

#include "hipdvr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sum of 8 Gaussian blobs (centres U[0.2,0.8]^3, sigma U[0.05,0.15],
 * amplitude U[0.3,1.0]) over normalized voxel centres, scaled so the maximum
 * maps to the type's range (255 / 65535), plus per-voxel uniform noise of
 * +-2 steps (u8) or +-2*257 (u16), clamped.  dtype: 0 = uint8, 1 = uint16.
 * out: x-fastest, nx*ny*nz voxels. */
int hipdvr_synth_blobs(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, int32_t dtype,
                       void* out);

/* Nested ellipsoid shells (semi-axes 0.62/0.56/0.50 of the unit cube) at 5
 * densities {0.15, 0.45, 0.70, 0.25, 0.90} from outside in, plus 4-octave fBm
 * value noise of amplitude 5% of the range; voxels outside the outer shell
 * (about 30% of the cube) are air, exactly 0.  uint16. */
int hipdvr_synth_shells(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, uint16_t* out);

/* Turbulence on the device, for volumes too large to make on the host: the
 * sum over 6 octaves of 2^-o |value noise| (frequency 6*2^o lattice cells
 * across the largest axis), normalized and quantized to uint16, then
 * thresholded so that the fraction of voxels nearest `empty_fraction` that a
 * boundary of a 4096-bin histogram allows is exactly 0 and the rest is
 * rescaled linearly onto [1, 65535] (the top of the highest occupied bin
 * maps to 65535).  d_out: device memory, x-fastest nx*ny*nz uint16.  The exact
 * fraction of zero voxels is stored in *empty_out (may be NULL).
 * Synchronous on the context stream. */
int hipdvr_synth_turbulence_device(hipdvr_ctx* ctx, uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz,
                                   float empty_fraction, uint16_t* d_out, double* empty_out);

/* The same volume as hipdvr_synth_turbulence_device, bit for bit, without
 * allocating: no hipMalloc, hipFree or other allocating call.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under stream capture, serving loops, harnesses that forbid it).  The
 * generator's scratch -- the 4096-bin histogram and the zero counter, 32,776
 * bytes whatever the volume size -- is a file-scope device array of the
 * library, one copy per GPU, held until the process ends; its address on a
 * device is looked up once by hipdvr_create on that device.  Every call
 * clears it first (with a fill kernel).  Calls on one GPU share it: calls
 * that overlap on one GPU (several contexts or host threads) race and must
 * be serialized by the caller. */
int hipdvr_synth_turbulence_device_noalloc(hipdvr_ctx* ctx, uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz,
                                           float empty_fraction, uint16_t* d_out, double* empty_out);

#ifdef __cplusplus
}
#endif
#endif
