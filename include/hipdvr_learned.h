/*
 * hipdvr_learned.h -- learned stepping for the DVR ray march.
 *
 * No counterpart in the reference: Voreen's raycasters march at a fixed step
 * (cpuraycaster.cpp:104,315) and skip nothing.  This is the learned addition
 * beside the drop-in path.  A small network looks at per-brick statistics of
 * the volume and at the transfer function, and proposes for each 16^3 brick
 * and ray-direction octant:
 *   - whether the brick can be skipped;
 *   - a step multiplier (1, 2, 4 or 8 times the plain step).
 *
 * The network has 16 inputs, two hidden layers of 64 (GELU) and 2 outputs.
 * Its weights are BF16 with FP32 accumulation (MFMA). It is trained on the GPU
 * by the library's own forward/backward/Adam kernels, on targets measured
 * from plain-marcher traces.
 *
 * The plain marcher stays the judge. For each learned frame, 1/64 of the rays
 * are re-marched plainly and compared: 8x8-pixel blocks, one block in every
 * 64x64 pixels (hipdvr_learned_sparse_dims gives the grid).  If the PSNR of
 * that sparse set falls below the floor (45 dB by default), the frame is
 * rendered plainly instead.
 *
 * Features of one (brick, octant) row, normalized intensities in [0, 1]:
 *   0 min  1 max  2 mean  3 std of the brick's voxels (16^3 + 1-voxel apron)
 *   4 mean corrected TF alpha over the index range [idx(min), idx(max)]
 *   5 corrected TF alpha at idx(mean)
 *   6 fraction of that range with non-zero alpha
 *   7 opacity of 32 samples at the mean alpha: 1 - (1 - f4)^32
 *   8..10 ray direction octant signs (+1 / -1)
 *   11..15 zero
 * Targets (3 floats): skip label in {0,1}, log2 multiplier label in [0,3],
 * multiplier-loss weight.
 */
#ifndef HIPDVR_LEARNED_H
#define HIPDVR_LEARNED_H

#include <stdint.h>

#include "hipdvr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HIPDVR_MLP_INPUTS 16
#define HIPDVR_MLP_OUTPUTS 2
#define HIPDVR_MLP_TARGETS 3
/* W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2] */
#define HIPDVR_MLP_PARAMS (16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2)

typedef struct hipdvr_mlp hipdvr_mlp;

/* network with seeded Xavier-uniform weights and zero biases */
int hipdvr_mlp_create(hipdvr_ctx* ctx, uint64_t seed, hipdvr_mlp** out);
int hipdvr_mlp_destroy(hipdvr_mlp* mlp);
/* FP32 master parameters (HIPDVR_MLP_PARAMS floats), host arrays */
int hipdvr_mlp_get_params(hipdvr_mlp* mlp, float* out);
int hipdvr_mlp_set_params(hipdvr_mlp* mlp, const float* in);
/* device arrays: x [rows][16] -> y [rows][2] */
int hipdvr_mlp_forward(hipdvr_mlp* mlp, const float* d_x, uint32_t rows, float* d_y);
/* gradient of the mean loss over the rows (host output: HIPDVR_MLP_PARAMS
 * floats) and the loss itself; device inputs x [rows][16], t [rows][3] */
int hipdvr_mlp_gradient(hipdvr_mlp* mlp, const float* d_x, const float* d_t, uint32_t rows,
                        float* grad_out, float* loss_out);
/* one Adam step on a minibatch (device arrays); loss_out may be NULL */
int hipdvr_mlp_train_step(hipdvr_mlp* mlp, const float* d_x, const float* d_t, uint32_t rows,
                          float lr, float* loss_out);

/* ---- learned rendering ---- */
typedef struct hipdvr_learned hipdvr_learned;

typedef struct hipdvr_learned_opts {
    float psnr_floor;        /* dB; below it the frame is rendered plainly (default 45) */
    int32_t remarch_stride;  /* sparse judge: 1/n of the pixels along x and y, in blocks of 8 (default 8 = 1/64) */
    float skip_logit;        /* skip a brick when the skip logit exceeds this (default 0) */
    int32_t max_log2_step;   /* largest step multiplier, as log2 (0..3, default 3) */
} hipdvr_learned_opts;

typedef struct hipdvr_learned_stats {
    uint64_t samples;        /* samples the learned marcher reconstructed */
    uint64_t shaded;
    float sparse_psnr;       /* PSNR of the learned image at the sparse pixels vs plain */
    int32_t fell_back;       /* 1 if the frame was rendered plainly */
} hipdvr_learned_stats;

/* Learned state of one volume: 16^3 brick statistics, features, the
 * (brick, octant) stepping table, trace accumulators. */
int hipdvr_learned_create(hipdvr_ctx* ctx, const hipdvr_volume* vol, hipdvr_learned** out);
int hipdvr_learned_destroy(hipdvr_learned* l);
/* Features for this TF and step; with mlp != NULL the table comes from the
 * network, with mlp == NULL it is the no-network baseline (multiplier 1, skip
 * only bricks whose value range maps to zero alpha). */
int hipdvr_learned_prepare(hipdvr_learned* l, const hipdvr_tf* tf, const hipdvr_params* prm,
                           hipdvr_mlp* mlp, const hipdvr_learned_opts* opts);
/* Training data: clear the trace accumulators, trace one view (device
 * entry/exit) into them, and append the visited rows (features of the last
 * prepare + targets) to device arrays x [capacity][16], t [capacity][3]
 * starting at *count (updated).  Tracing needs a volume below 4 GiB
 * (HIPDVR_ERR_UNSUPPORTED otherwise); a network trained on a smaller volume
 * of the same kind serves a larger one, since it reads brick statistics. */
int hipdvr_learned_trace_reset(hipdvr_learned* l);
int hipdvr_learned_trace(hipdvr_learned* l, const hipdvr_tf* tf, const hipdvr_params* prm,
                         const float* d_entry, const float* d_exit, uint32_t w, uint32_t h);
int hipdvr_learned_collect(hipdvr_learned* l, float tau_skip, float tau_step, float* d_x, float* d_t,
                           uint32_t capacity, uint32_t* count);
/* Training of the network for one volume, the whole loop on the device:
 * for each of n_tfs seeded random transfer functions (4-8 keys, intensities
 * U[0,1], RGBA8 U{0..255}) and each of the caller's extra tables, trace the
 * volume plainly from n_views orbit cameras (radius 3.5 around the proxy box,
 * fovy 30, elevations cycling 10..60 degrees) at trace_size^2, collect the
 * visited rows with targets at tau_skip / tau_step, then take
 * max(min_steps, epochs * rows / batch) Adam steps at rate lr on random
 * contiguous minibatches.  Zero fields take the defaults in brackets.
 * Replaces the learned state's last prepare (call hipdvr_learned_prepare
 * again before rendering).  Synchronous. */
typedef struct hipdvr_learned_train_opts {
    uint32_t n_tfs;          /* random TFs traced [200] */
    uint32_t n_views;        /* orbit views per TF [16] */
    uint32_t trace_size;     /* trace image side in pixels [192] */
    float sampling_rate;     /* step 1/(max(dim)*rate) [2] */
    float tau_skip;          /* skip-label error threshold [3e-3] */
    float tau_step;          /* step-label error threshold [8e-3] */
    uint32_t min_steps;      /* Adam steps at least [2000] */
    uint32_t epochs;         /* passes over the rows at most [4] */
    uint32_t batch;          /* minibatch rows [8192] */
    float lr;                /* Adam learning rate [2e-3] */
    uint64_t seed;           /* TF and minibatch seed [0x5EED] */
    const uint8_t* extra_tfs; /* n_extra RGBA8 tables of extra_width entries, traced after the random ones */
    uint32_t n_extra;
    uint32_t extra_width;    /* [256 for uint8 volumes, 4096 for uint16] */
    uint64_t max_rows;       /* training rows at most [24,000,000] */
} hipdvr_learned_train_opts;

typedef struct hipdvr_learned_train_info {
    uint64_t rows;           /* training rows collected */
    uint32_t tfs;            /* tables traced */
    uint32_t views;
    uint32_t trace_size;
    uint32_t steps;          /* Adam steps taken */
    float loss_first, loss_last;
    float trace_seconds;     /* tracing and collecting */
    float seconds;           /* the whole call */
} hipdvr_learned_train_info;

int hipdvr_learned_train(hipdvr_learned* l, hipdvr_mlp* mlp, const hipdvr_learned_train_opts* opts,
                         hipdvr_learned_train_info* info);

/* Render with learned stepping; the sparse plain re-march judges it and the
 * frame falls back to the plain marcher below the PSNR floor.  When the last
 * prepared table keeps every brick that may hold alpha > 0 at the base step
 * (no gain possible), the plain frame is rendered directly, without the
 * judge (stats: sparse_psnr 999, fell_back 0).  Asynchronous unless
 * stats != NULL. */
int hipdvr_render_learned_device(hipdvr_learned* l, const hipdvr_tf* tf, const float* d_entry,
                                 const float* d_exit, uint32_t w, uint32_t h, const hipdvr_params* prm,
                                 const hipdvr_learned_opts* opts, float* d_out,
                                 hipdvr_learned_stats* stats);
/* The judge's sparse grid for a w x h image: along each axis blocks of 8
 * consecutive pixels, one block in every 8 * remarch_stride pixels, starting
 * at (remarch_stride - 1) * 4 clamped to the image (grid column g -> pixel
 * (g / 8) * 8 * stride + offset + g % 8).  *sw x *sh is the number of grid
 * pixels; a ragged image edge can add up to one partial block per axis, so
 * it may exceed ceil(w / stride) x ceil(h / stride).  Host arithmetic only
 * (opts may be NULL: stride 8). */
int hipdvr_learned_sparse_dims(uint32_t w, uint32_t h, const hipdvr_learned_opts* opts, uint32_t* sw,
                               uint32_t* sh);
/* The judge's plain re-march on its own, on the grid of
 * hipdvr_learned_sparse_dims, marched by the plain path with each ray cut
 * into segments that run in parallel and are joined front to back (equal to
 * the sequential march up to float rounding of the joins).  d_sparse holds
 * `capacity` RGBA float pixels and receives *sw x *sh of them (row-major);
 * HIPDVR_ERR_INVALID when capacity is below that.  Synchronous. */
int hipdvr_learned_sparse_remarch(hipdvr_learned* l, const hipdvr_tf* tf, const float* d_entry,
                                  const float* d_exit, uint32_t w, uint32_t h, const hipdvr_params* prm,
                                  const hipdvr_learned_opts* opts, float* d_sparse, uint64_t capacity,
                                  uint32_t* sw, uint32_t* sh);

#ifdef __cplusplus
}
#endif
#endif
