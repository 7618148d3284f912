/*
 * hipdvr_io.h -- volume files and whole-volume statistics.
 *
 * hipdvr_vvd_parse / hipdvr_volume_load_vvd replace the host-side reader
 * VvdVolumeReader (voreen/modules/core/io/vvdvolumereader.cpp, format
 * voreen/modules/core/io/vvdformat.cpp:53-90): the .vvd XML names the .raw
 * file ("Paths" item list, or the older "filename" attribute relative to the
 * document), the voxel format and the dimensions; meta items carry Spacing,
 * Offset (Vec3MetaData) and RealWorldMapping (scale, offset;
 * src/core/datastructures/meta/realworldmappingmetadata.cpp:53-62).  Only the
 * first volume of a file is read; uint8 and uint16 voxels are supported.
 *
 * hipdvr_volume_minmax replaces the derived data VolumeMinMax
 * (include/voreen/core/datastructures/volume/volumeminmax.h) that a TF
 * domain is fitted to: the normalized minimum and maximum voxel
 * (VolumeAtomic::getVoxelNormalized, volumeatomic.h:531-536) and their
 * real-world values, reduced on the device from the brick min/max table.
 */
#ifndef HIPDVR_IO_H
#define HIPDVR_IO_H

#include <stdint.h>

#include "hipdvr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host only (no device needed): fill desc (dims, dtype, rwm, spacing,
 * offset; brick_log2 = 0) and the absolute or vvd-relative raw file path
 * (NUL-terminated, at most raw_path_cap bytes incl. the NUL). */
int hipdvr_vvd_parse(const char* vvd_path, hipdvr_volume_desc* desc, char* raw_path, uint32_t raw_path_cap);
/* message of the last failed hipdvr_vvd_parse on this thread */
const char* hipdvr_vvd_last_error(void);

/* Read dims[0]*dims[1]*dims[2] voxels of desc->dtype from byte_offset of a
 * raw file (x-fastest), streamed through pinned memory, and build the device
 * volume as hipdvr_volume_create does. */
int hipdvr_volume_load_raw(hipdvr_ctx* ctx, const char* raw_path, const hipdvr_volume_desc* desc,
                           uint64_t byte_offset, hipdvr_volume** out);
/* Parse + load; desc_out (may be NULL) receives the parsed description. */
int hipdvr_volume_load_vvd(hipdvr_ctx* ctx, const char* vvd_path, uint32_t brick_log2, hipdvr_volume** out,
                           hipdvr_volume_desc* desc_out);

/* out[0] minNorm, out[1] maxNorm, out[2] min, out[3] max (real world) */
int hipdvr_volume_minmax(hipdvr_volume* vol, float out[4]);

#ifdef __cplusplus
}
#endif
#endif
