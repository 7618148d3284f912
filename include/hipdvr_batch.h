/*
 * hipdvr_batch.h -- many volumes, each with its own TF and camera, marched in
 * one persistent launch.
 *
 * No counterpart in the reference: Voreen renders one volume per raycaster
 * process() call (cpuraycaster.cpp:90-208; a network of N raycasters makes N
 * calls).  A batch is the library's answer for scenes of many uneven volumes
 * (multi-view dashboards, ensembles): every job is the plain DVR march of
 * hipdvr_render_device (linear filter, TF classification, DVR compositing,
 * optional Phong) with the same result bit for bit, but all jobs' image tiles
 * are handed out by one device work queue to a persistent grid of workgroups,
 * largest jobs first.  The launch can be captured into a hipGraph and
 * replayed; an autotuner picks the wave tile shape, the empty-space brick
 * size and the number of samples fetched together per ray.
 */
#ifndef HIPDVR_BATCH_H
#define HIPDVR_BATCH_H

#include <stdint.h>

#include "hipdvr.h"
#include "hipdvr_learned.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hipdvr_batch hipdvr_batch;

typedef struct hipdvr_batch_job {
    const hipdvr_volume* volume;
    const hipdvr_tf* tf;
    const float* d_entry;   /* device RGBA32F w*h (hipdvr_entry_exit_device) */
    const float* d_exit;
    float* d_out;           /* device RGBA32F w*h */
    uint32_t w, h;
    hipdvr_params params;   /* step, shading, light, camera; filter must be linear,
                               classification TF, compositing DVR */
} hipdvr_batch_job;

/* wave tile shapes (pixels per wavefront) */
enum { HIPDVR_TILE_8x8 = 0, HIPDVR_TILE_16x4 = 1, HIPDVR_TILE_32x2 = 2 };

typedef struct hipdvr_batch_config {
    int32_t tile;           /* HIPDVR_TILE_* */
    int32_t brick_log2;     /* empty-space bricks of 8^3, 16^3, 32^3: 3, 4, 5 */
    int32_t batch;          /* samples fetched together per ray: 2 or 4 */
    int32_t resident;       /* persistent workgroups per CU, 1..8 (at most what fits); 0: three
                               when a job's cell records need 64-bit addressing (their caches
                               thrash sooner), else as many as fit, at most five */
} hipdvr_batch_config;

/* Jobs are copied; the volumes, TFs and buffers must outlive the batch.
 * Builds each volume's cell records and, per candidate brick size, each
 * job's occupancy table for its TF. */
int hipdvr_batch_create(hipdvr_ctx* ctx, const hipdvr_batch_job* jobs, uint32_t n, hipdvr_batch** out);
int hipdvr_batch_destroy(hipdvr_batch* b);
/* select a configuration (drops a captured graph) */
int hipdvr_batch_set_config(hipdvr_batch* b, const hipdvr_batch_config* cfg);
int hipdvr_batch_get_config(const hipdvr_batch* b, hipdvr_batch_config* cfg);
/* Time every tile shape, brick size and batch over reps launches each (HIP
 * events) at the default residency, then 3, 4 and 5 resident workgroups per
 * CU with the fastest; keep the fastest; best_ms (may be NULL) receives its
 * mean launch time. */
int hipdvr_batch_autotune(hipdvr_batch* b, uint32_t reps, hipdvr_batch_config* best, float* best_ms);
/* Capture one batched launch (queue reset + persistent march) as a hipGraph
 * on the context stream; hipdvr_batch_launch then replays it. */
int hipdvr_batch_capture(hipdvr_batch* b);
/* Asynchronous on the context stream: the graph when captured, else a direct
 * launch; with stats != NULL it is a counted, synchronous launch. */
int hipdvr_batch_launch(hipdvr_batch* b, hipdvr_stats* stats);

/* Learned stepping in the batch: learned[i] (in the order the jobs were given
 * to hipdvr_batch_create; NULL keeps job i plain) is a learned state of job
 * i's volume whose last hipdvr_learned_prepare was for job i's TF.  Used by
 * hipdvr_batch_launch_learned only (hipdvr_batch_launch and a captured graph
 * stay plain).  Jobs sharing a learned state share its step; jobs shading
 * with the exact Phong path cannot be learned.  A job whose state's table
 * cannot gain (every brick that may hold alpha > 0 at the base step) is
 * marched plainly (stats: sparse_psnr 999).  A later prepare of a state
 * (a new network, the same TF) takes effect at the next launch; the states
 * must outlive their use by the batch.  Calling again replaces the set. */
int hipdvr_batch_set_learned(hipdvr_batch* b, hipdvr_learned* const* learned, uint32_t n,
                             const hipdvr_learned_opts* opts);
/* One batched launch with learned stepping: one persistent launch marches
 * the judges' plain re-march of every learned view's sparse pixels (each ray
 * whole, so the sparse images equal the plain frames at those pixels
 * exactly) and the learned and plain jobs; then the judges of all learned
 * views run in one launch, and per learned view the plain frame, which runs
 * only when its view fell below the PSNR floor (the steps of
 * hipdvr_render_learned_device, for all views of the batch at once; that
 * entry point re-marches one view's judge rays in segments, equal up to
 * float rounding).  Asynchronous unless stats != NULL:
 * n_jobs entries in the caller's order, sparse_psnr and fell_back per learned
 * job, samples and shaded the totals of the whole launch. */
int hipdvr_batch_launch_learned(hipdvr_batch* b, hipdvr_learned_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
