// batch.hip -- many volumes in one persistent launch (include/hipdvr_batch.h).
//
// Every job is the cell-record march of hipdvr_render_device (march_cell.h),
// so each job's image is bit-identical to its single-frame render.  The image
// tiles of all jobs form one list, jobs in decreasing order of estimated cost
// (pixels x longest volume axis: largest first); a persistent grid of
// workgroups (CUs x resident workgroups per CU, at most five) takes tiles
// from a device counter until the list is exhausted, reloading the TF and
// opacity table in LDS when a tile belongs to another job.  Every workgroup leaves when the
// counter passes the end of the list, so the grid always drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/hipdvr_batch.h"
#include "../../include/hipdvr_learned.h"
#include "ctx.h"
#include "errors.h"
#include "kernels.h"
#include "learned.h"
#include "learned_ray.h"
#include "march_cell.h"
#include "objects.h"
#include "volume.h"

namespace hipdvr {

namespace {

constexpr int kTileShapes = 3;
constexpr int kMaxResidentPerCu = 5;
constexpr int kBrickSizes = 3;   // brick_log2 3, 4, 5

// pixels per wavefront: 8x8, 16x4, 32x2; a workgroup is 2x2 wavefronts
__host__ __device__ constexpr int wave_w(int tile) { return tile == 0 ? 8 : (tile == 1 ? 16 : 32); }
__host__ __device__ constexpr int wave_h(int tile) { return 64 / wave_w(tile); }

// Tile queues.  The hardware hands the persistent grid's workgroups to the
// 8 XCDs in turn (workgroup b to XCD b % 8), and each XCD has its own L2.
// With kXcdQueues the tiles come in blocks of kBlockW x kBlockH tiles of one
// job (blocks row by row, tiles row by row inside a block; edge blocks
// smaller); consecutive runs of kChunk tiles of the launch's tile range are
// dealt to the XCDs in turn, and each XCD takes its runs from a counter of
// its own, so the tiles an XCD marches at a time form a compact column of
// blocks whose rays share that XCD's L2.  An XCD whose runs are taken takes
// from the others' counters.  Without it: one counter, tiles row by row.
constexpr bool kXcdQueues = true;
constexpr int kBlockW = 8, kBlockH = 4;
constexpr uint32_t kChunk = kBlockW * kBlockH;
constexpr uint32_t kCounterStride = 32;   // counters 128 bytes apart

__global__ void batch_reset_kernel(uint32_t* counter) {
    if (threadIdx.x < kXcds) counter[threadIdx.x * kCounterStride] = 0;
}

// Tiles dealt to XCD x out of total.
__device__ __forceinline__ uint32_t xcd_tiles(uint32_t x, uint32_t total) {
    const uint32_t chunks = (total + kChunk - 1) / kChunk;
    if (x >= chunks) return 0;
    const uint32_t owned = (chunks - 1 - x) / kXcds + 1;
    uint32_t n = owned * kChunk;
    if ((chunks - 1) % kXcds == x) n -= chunks * kChunk - total;   // its last run is the short one
    return n;
}

// Tile (tx, ty) at position lt of a job of tiles_x x tiles_y tiles in block
// order (a bijection of [0, tiles_x * tiles_y)).
__device__ __forceinline__ void block_tile(uint32_t lt, uint32_t tiles_x, uint32_t tiles_y, uint32_t& tx,
                                           uint32_t& ty) {
    const uint32_t nby = (tiles_y + kBlockH - 1) / kBlockH, nbx = (tiles_x + kBlockW - 1) / kBlockW;
    uint32_t by = lt / (kBlockH * tiles_x);
    if (by > nby - 1) by = nby - 1;
    const uint32_t r = lt - by * kBlockH * tiles_x;
    const uint32_t bh = min((uint32_t)kBlockH, tiles_y - by * kBlockH);
    uint32_t bx = r / (kBlockW * bh);
    if (bx > nbx - 1) bx = nbx - 1;
    const uint32_t r2 = r - bx * kBlockW * bh;
    const uint32_t bw = min((uint32_t)kBlockW, tiles_x - bx * kBlockW);
    ty = by * kBlockH + r2 / bw;
    tx = bx * kBlockW + r2 % bw;
}

template <bool WIDE, int SHADE, bool COUNT, int K, int TILE>
__global__ __launch_bounds__(256) void batch_kernel(const LearnedArgs* __restrict__ jobs,
                                                    const uint32_t* __restrict__ tile_start, uint32_t njobs,
                                                    uint32_t total_tiles, uint32_t* __restrict__ counter,
                                                    unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    __shared__ uint32_t s_tile;
    constexpr int WW = wave_w(TILE), WH = wave_h(TILE);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int cur_job = -1;
    RayCounters n;
    uint32_t drained = 0;   // (lane 0) XCD queues seen empty
    const uint32_t my_xcd = blockIdx.x % kXcds;
    while (true) {
        if (tid == 0) {
            if (kXcdQueues) {
                uint32_t t = total_tiles;
                for (uint32_t k = 0; k < kXcds && t == total_tiles; ++k) {
                    const uint32_t x = (my_xcd + k) % kXcds;
                    if (drained & (1u << x)) continue;
                    const uint32_t i = atomicAdd(counter + x * kCounterStride, 1u);
                    if (i < xcd_tiles(x, total_tiles)) t = (x + kXcds * (i / kChunk)) * kChunk + i % kChunk;
                    else drained |= 1u << x;
                }
                s_tile = t;
            } else {
                s_tile = atomicAdd(counter, 1u);
            }
        }
        __syncthreads();   // also: every wave has finished the previous tile
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();   // s_tile is read before the next iteration writes it
        if (tile >= total_tiles) break;
        // job of this tile: last j with tile_start[j] <= tile
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tile_start[mid] <= tile) lo = mid;
            else hi = mid;
        }
        const int job = (int)__builtin_amdgcn_readfirstlane(lo);
        const LearnedArgs& la = jobs[job];
        const MarchArgs& a = la.m;
        // a job with a stepping table marches with learned stepping
        const bool learned = la.table != nullptr;
        if (job != cur_job) {
            for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
            s_alpha[tid] = learned ? la.alpha_m[tid] : a.alpha_corr[tid];
            __syncthreads();
            cur_job = job;
        }
        // a job marches the pixels of its grid: the whole image, or (a
        // judge's re-march: pix_block > 1) the sparse pixels judge_pixel
        // places, written to out[gy * grid_w + gx]; their samples are not
        // counted (the learned frame's statistics leave the judge out)
        const bool sparse = a.pix_block > 1;
        RayCounters uncounted;
        RayCounters& nc = sparse ? uncounted : n;
        const uint32_t lt = tile - tile_start[job];
        const uint32_t tiles_x = (uint32_t)(a.grid_w + 2 * WW - 1) / (2 * WW);
        uint32_t tx = lt % tiles_x, ty = lt / tiles_x;
        if (kXcdQueues) block_tile(lt, tiles_x, (uint32_t)(a.grid_h + 2 * WH - 1) / (2 * WH), tx, ty);
        const int gx = (int)tx * (2 * WW) + (wave & 1) * WW + (lane % WW);
        const int gy = (int)ty * (2 * WH) + (wave >> 1) * WH + (lane / WW);
        if (gx < a.grid_w && gy < a.grid_h) {
            const int px = sparse ? judge_pixel(gx, a.stride, a.pix_block, a.pix_offset) : gx;
            const int py = sparse ? judge_pixel(gy, a.stride, a.pix_block, a.pix_offset) : gy;
            const size_t pix = (size_t)py * a.w + px;
            const float4 e = a.entry[pix];
            const float4 l = a.exit[pix];
            float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!is_background(e, l)) {
                // WIDE: 64-bit record addressing for every job (a volume's
                // records exceed 4 GiB); else 32-bit buffer offsets
                if (learned) {
                    unsigned long long ns = 0, nsh = 0;
                    if (a.dtype == 0) {
                        const CellFetch<uint8_t, WIDE> f = cell_fetch_of<uint8_t, WIDE>(a);
                        res = learned_ray<uint8_t, WIDE, SHADE != 0>(la, f, s_tf, s_alpha, px, py, e, l, ns, nsh);
                    } else {
                        const CellFetch<uint16_t, WIDE> f = cell_fetch_of<uint16_t, WIDE>(a);
                        res = learned_ray<uint16_t, WIDE, SHADE != 0>(la, f, s_tf, s_alpha, px, py, e, l, ns, nsh);
                    }
                    ++n.rays;
                    n.samples += ns;
                    n.fetched += ns;
                    n.shaded += nsh;
                } else if (a.dtype == 0) {
                    const CellFetch<uint8_t, WIDE> f = cell_fetch_of<uint8_t, WIDE>(a);
                    res = cell_ray<uint8_t, WIDE, SHADE, COUNT, K>(a, f, s_tf, s_alpha, nullptr, px, py, e, l, nc);
                } else {
                    const CellFetch<uint16_t, WIDE> f = cell_fetch_of<uint16_t, WIDE>(a);
                    res = cell_ray<uint16_t, WIDE, SHADE, COUNT, K>(a, f, s_tf, s_alpha, nullptr, px, py, e, l, nc);
                }
            }
            a.out[(size_t)gy * a.grid_w + gx] = res;
        }
    }
    if (COUNT) {
        wave_add(counters + 0, n.rays);
        wave_add(counters + 1, n.samples);
        wave_add(counters + 2, n.fetched);
        wave_add(counters + 3, n.shaded);
        wave_add_max64(counters + 4, n.slots);
        wave_add_max64(counters + 5, n.shade_slots);
    }
}

typedef void (*BatchKernel)(const LearnedArgs*, const uint32_t*, uint32_t, uint32_t, uint32_t*,
                            unsigned long long*);

template <bool WIDE, int SHADE, bool COUNT>
BatchKernel pick_kernel(int tile, int k) {
    if (k <= 2) {
        if (tile == 0) return batch_kernel<WIDE, SHADE, COUNT, 2, 0>;
        if (tile == 1) return batch_kernel<WIDE, SHADE, COUNT, 2, 1>;
        return batch_kernel<WIDE, SHADE, COUNT, 2, 2>;
    }
    if (tile == 0) return batch_kernel<WIDE, SHADE, COUNT, 4, 0>;
    if (tile == 1) return batch_kernel<WIDE, SHADE, COUNT, 4, 1>;
    return batch_kernel<WIDE, SHADE, COUNT, 4, 2>;
}

template <bool WIDE>
BatchKernel kernel_for_w(int shade, bool count, int tile, int k) {
    if (shade == 2) return count ? pick_kernel<WIDE, 2, true>(tile, k) : pick_kernel<WIDE, 2, false>(tile, k);
    if (shade == 1) return count ? pick_kernel<WIDE, 1, true>(tile, k) : pick_kernel<WIDE, 1, false>(tile, k);
    return count ? pick_kernel<WIDE, 0, true>(tile, k) : pick_kernel<WIDE, 0, false>(tile, k);
}

BatchKernel kernel_for(bool wide, int shade, bool count, int tile, int k) {
    return wide ? kernel_for_w<true>(shade, count, tile, k) : kernel_for_w<false>(shade, count, tile, k);
}

}  // namespace

}  // namespace hipdvr

using namespace hipdvr;

struct hipdvr_batch {
    hipdvr_ctx* ctx = nullptr;
    std::vector<hipdvr_batch_job> jobs;    // in launch order (largest first)
    std::vector<uint32_t> order;           // caller's index of each launch position
    int shade = 0;                         // 0 none, 1 Phong, 2 exact Phong
    bool wide = false;                     // some volume's cell records exceed 4 GiB
    int sms = 0;
    // per brick size: device brick tables per job and the jobs' arguments
    // (march arguments, plus the stepping table of learned jobs)
    std::vector<void*> allocations;
    std::vector<LearnedArgs> h_args[kBrickSizes];
    LearnedArgs* d_args[kBrickSizes] = {nullptr, nullptr, nullptr};
    // the same with the stepping tables of the learned jobs
    // (hipdvr_batch_launch_learned; hipdvr_batch_launch stays plain)
    LearnedArgs* d_largs[kBrickSizes] = {nullptr, nullptr, nullptr};
    // learned jobs (hipdvr_batch_set_learned), per launch position: the judge's
    // grid, sparse image, sum and flag
    std::vector<hipdvr_learned*> learned;
    std::vector<char> plain_learned;   // given a learned state whose table cannot gain: marched plainly
    hipdvr_learned_opts lopts{};
    std::vector<hipdvr::JudgeGrid> grids;
    std::vector<float4*> d_sparse;
    std::vector<float*> d_sse;
    std::vector<int*> d_flag;
    // the learned launch's job list: first each learned view's judge
    // re-march (its sparse grid, marched plainly into its sparse image),
    // then every job; its tile ranges per tile shape
    uint32_t n_ljobs = 0;
    uint32_t* d_ltile_start[kTileShapes] = {nullptr, nullptr, nullptr};
    uint32_t ltotal_tiles[kTileShapes] = {0, 0, 0};
    hipdvr::JudgeJob* d_judge = nullptr;
    uint32_t n_judge = 0;
    uint32_t* d_tile_start[kTileShapes] = {nullptr, nullptr, nullptr};
    uint32_t total_tiles[kTileShapes] = {0, 0, 0};
    uint32_t* d_counter = nullptr;
    hipdvr_batch_config cfg{0, 4, 4, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(hipdvr_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_batch(hipdvr_batch* b) {
    if (!b) return;
    drop_graph(b);
    for (void* p : b->allocations) (void)hipFree(p);
    delete b;
}

template <typename P>
P* dev_alloc(hipdvr_batch* b, size_t bytes) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 4), "hipMalloc(batch)");
    b->allocations.push_back(p);
    return (P*)p;
}

void validate_config(const hipdvr_batch_config& c) {
    if (c.tile < HIPDVR_TILE_8x8 || c.tile > HIPDVR_TILE_32x2) fail(HIPDVR_ERR_INVALID, "unknown tile shape");
    if (c.brick_log2 < 3 || c.brick_log2 > 5) fail(HIPDVR_ERR_INVALID, "brick_log2 must be 3, 4 or 5");
    if (c.batch != 2 && c.batch != 4) fail(HIPDVR_ERR_INVALID, "batch must be 2 or 4");
    if (c.resident < 0 || c.resident > 8) fail(HIPDVR_ERR_INVALID, "resident must be 0 (default) or 1..8");
}

void launch_batch(hipdvr_batch* b, bool count, bool learned = false) {
    hipStream_t st = ctx_stream(b->ctx);
    const int t = b->cfg.tile, bs = b->cfg.brick_log2 - 3;
    BatchKernel k = kernel_for(b->wide, b->shade, count, t, b->cfg.batch);
    int per_cu = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, 256, 0), "occupancy query");
    if (per_cu < 1) per_cu = 1;
    // resident workgroups per CU: more rays in flight evict each other's cell
    // records.  Default: at most five (64 uneven volumes, unshaded: 59,030
    // frames/s at the occupancy limit, 63,436 at five, 62,716 at four), three
    // when records need 64-bit addressing (8 GiB volume, 8 views of 4K with
    // Phong: 234 ms at three against 277 ms at four)
    const int cap = b->cfg.resident > 0 ? b->cfg.resident : (b->wide ? 3 : kMaxResidentPerCu);
    if (per_cu > cap) per_cu = cap;
    const uint32_t total = learned ? b->ltotal_tiles[t] : b->total_tiles[t];
    uint32_t grid = (uint32_t)(b->sms * per_cu);
    if (grid > total) grid = total ? total : 1;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, b->d_counter);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st,
                       (const LearnedArgs*)(learned ? b->d_largs[bs] : b->d_args[bs]),
                       (const uint32_t*)(learned ? b->d_ltile_start[t] : b->d_tile_start[t]),
                       learned ? b->n_ljobs : (uint32_t)b->jobs.size(), total, b->d_counter, b->ctx->d_counters);
    hip_check(hipGetLastError(), "batch launch");
}

}  // namespace

extern "C" {

int hipdvr_batch_create(hipdvr_ctx* ctx, const hipdvr_batch_job* jobs, uint32_t n, hipdvr_batch** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        if (!out || !jobs || n == 0) fail(HIPDVR_ERR_INVALID, "a batch needs at least one job");
        int shade = -1;
        for (uint32_t i = 0; i < n; ++i) {
            const hipdvr_batch_job& j = jobs[i];
            if (!j.volume || !j.tf || !j.d_entry || !j.d_exit || !j.d_out || !j.w || !j.h)
                fail(HIPDVR_ERR_INVALID, "job " + std::to_string(i) + ": NULL argument or empty image");
            if (j.volume->ctx != ctx || j.tf->ctx != ctx) fail(HIPDVR_ERR_INVALID, "handle from another context");
            validate_params(&j.params, j.tf);
            if (j.params.filter != HIPDVR_FILTER_LINEAR || j.params.classification != HIPDVR_CLASS_TF ||
                j.params.compositing != HIPDVR_COMPOSITE_DVR)
                fail(HIPDVR_ERR_UNSUPPORTED, "batched jobs use the linear filter, TF classification and DVR");
            if (j.params.shading != HIPDVR_SHADE_NONE && j.params.gradient != HIPDVR_GRADIENT_CENTRAL)
                fail(HIPDVR_ERR_UNSUPPORTED, "batched jobs shade with central-difference gradients");
            const int ph = j.params.shading == HIPDVR_SHADE_NONE ? 0
                         : (j.params.shading == HIPDVR_SHADE_PHONG && !j.params.exact_shading) ? 1 : 2;
            if (shade >= 0 && ph != shade) fail(HIPDVR_ERR_UNSUPPORTED, "all jobs of a batch share the shading path");
            shade = ph;
        }
        ctx_set_device(ctx);
        hipStream_t st = ctx_stream(ctx);
        hipdvr_batch* b = new hipdvr_batch();
        b->ctx = ctx;
        b->shade = shade;
        try {
            int dev = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hip_check(hipDeviceGetAttribute(&b->sms, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
            // largest estimated cost first
            std::vector<uint32_t> order(n);
            std::iota(order.begin(), order.end(), 0u);
            auto cost = [&](uint32_t i) {
                const hipdvr_volume* v = jobs[i].volume;
                const uint32_t m = std::max(v->desc.dims[0], std::max(v->desc.dims[1], v->desc.dims[2]));
                return (double)jobs[i].w * jobs[i].h * m;
            };
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
            for (uint32_t i : order) b->jobs.push_back(jobs[i]);
            b->order = order;
            // cell records, then per brick size the brick tables of every job
            for (const hipdvr_batch_job& j : b->jobs) {
                ensure_cells(const_cast<hipdvr_volume*>(j.volume), st);
                b->wide = b->wide || j.volume->cells_wide;
            }
            for (int bs = 0; bs < kBrickSizes; ++bs) {
                std::vector<LearnedArgs>& args = b->h_args[bs];
                args.assign(n, LearnedArgs{});
                const uint32_t bl = (uint32_t)bs + 3;
                for (uint32_t i = 0; i < n; ++i) {
                    const hipdvr_batch_job& j = b->jobs[i];
                    const hipdvr_volume* v = j.volume;
                    uint32_t nb[3];
                    uint64_t count = 1;
                    for (int k = 0; k < 3; ++k) {
                        nb[k] = (v->desc.dims[k] + (1u << bl) - 1) >> bl;
                        count *= nb[k];
                    }
                    uint2* mm = dev_alloc<uint2>(b, count * sizeof(uint2));
                    uint8_t* occ = dev_alloc<uint8_t>(b, count);
                    const uint32_t words = (uint32_t)((count + 31) / 32);
                    uint32_t* bits = dev_alloc<uint32_t>(b, (size_t)words * 4);
                    launch_brick_minmax(v, bl, nb, mm, st);
                    OccArgs o;
                    o.minmax = mm;
                    o.occ = occ;
                    o.count = count;
                    o.dtype = v->desc.dtype;
                    o.rwm_scale = v->desc.rwm_scale;
                    o.rwm_offset = v->desc.rwm_offset;
                    o.dom_lo = j.tf->domain_lo;
                    o.dom_hi = j.tf->domain_hi;
                    o.tf_width = (int32_t)j.tf->width;
                    o.alpha_prefix = j.tf->d_alpha_prefix;
                    o.bits = bits;
                    launch_occupancy(o, st);
                    MarchArgs& a = args[i].m;
                    fill_march_args(a, v, j.tf, &j.params, j.d_entry, j.d_exit, j.d_out, j.w, j.h, nullptr);
                    a.cells = v->d_cells;
                    a.brick_log2 = bl;
                    for (int k = 0; k < 3; ++k) a.nbricks[k] = nb[k];
                    a.brick_occ = occ;
                    a.occ_bits = bits;
                    a.occ_words = words;
                }
                b->d_args[bs] = dev_alloc<LearnedArgs>(b, n * sizeof(LearnedArgs));
                hip_check(hipMemcpyAsync(b->d_args[bs], args.data(), n * sizeof(LearnedArgs), hipMemcpyHostToDevice,
                                         st),
                          "upload");
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            }
            for (int t = 0; t < kTileShapes; ++t) {
                std::vector<uint32_t> start(n + 1, 0);
                const uint32_t tw = 2 * wave_w(t), th = 2 * wave_h(t);
                for (uint32_t i = 0; i < n; ++i) {
                    const uint64_t tiles = (uint64_t)((b->jobs[i].w + tw - 1) / tw) * ((b->jobs[i].h + th - 1) / th);
                    if (start[i] + tiles > 0xfffffff0ull) fail(HIPDVR_ERR_UNSUPPORTED, "batch too large");
                    start[i + 1] = start[i] + (uint32_t)tiles;
                }
                b->total_tiles[t] = start[n];
                b->d_tile_start[t] = dev_alloc<uint32_t>(b, (n + 1) * 4);
                hip_check(hipMemcpyAsync(b->d_tile_start[t], start.data(), (n + 1) * 4, hipMemcpyHostToDevice, st),
                          "upload");
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            }
            b->d_counter = dev_alloc<uint32_t>(b, kXcds * kCounterStride * 4);
            b->learned.assign(n, nullptr);
            b->plain_learned.assign(n, 0);
            hip_check(hipGetLastError(), "batch setup");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            free_batch(b);
            throw;
        }
        *out = b;
    });
}

int hipdvr_batch_destroy(hipdvr_batch* b) {
    if (!b) return HIPDVR_OK;
    (void)hipStreamSynchronize(ctx_stream(b->ctx));
    free_batch(b);
    return HIPDVR_OK;
}

int hipdvr_batch_set_config(hipdvr_batch* b, const hipdvr_batch_config* cfg) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        if (!cfg) fail(HIPDVR_ERR_INVALID, "NULL argument");
        validate_config(*cfg);
        b->cfg = *cfg;
        drop_graph(b);
    });
}

int hipdvr_batch_get_config(const hipdvr_batch* b, hipdvr_batch_config* cfg) {
    if (!b || !cfg) return HIPDVR_ERR_INVALID;
    *cfg = b->cfg;
    return HIPDVR_OK;
}

int hipdvr_batch_capture(hipdvr_batch* b) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        ctx_set_device(b->ctx);
        hipStream_t st = ctx_stream(b->ctx);
        drop_graph(b);
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch_batch(b, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(st, &b->graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    });
}

int hipdvr_batch_launch(hipdvr_batch* b, hipdvr_stats* stats) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        ctx_set_device(b->ctx);
        hipStream_t st = ctx_stream(b->ctx);
        if (stats) {
            launch_zero_u64(b->ctx->d_counters, kNumCounters, st);
            hip_check(hipEventRecord(b->ctx->ev0, st), "hipEventRecord");
            launch_batch(b, true);
            hip_check(hipEventRecord(b->ctx->ev1, st), "hipEventRecord");
            unsigned long long c[kNumCounters];
            hip_check(hipMemcpyAsync(c, b->ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st), "download");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1), "hipEventElapsedTime");
            stats->rays = c[0];
            stats->samples = c[1];
            stats->fetched = c[2];
            stats->shaded = c[3];
            stats->lane_slots = c[4];
            stats->shade_slots = c[5];
            stats->kernel_ms = ms;
            return;
        }
        const bool timed = ctx_timing_start(b->ctx);
        if (b->exec) hip_check(hipGraphLaunch(b->exec, st), "hipGraphLaunch");
        else launch_batch(b, false);
        if (timed) ctx_timing_stop(b->ctx);
    });
}

int hipdvr_batch_autotune(hipdvr_batch* b, uint32_t reps, hipdvr_batch_config* best, float* best_ms) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        if (reps == 0) fail(HIPDVR_ERR_INVALID, "reps must be > 0");
        ctx_set_device(b->ctx);
        hipStream_t st = ctx_stream(b->ctx);
        drop_graph(b);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hipdvr_batch_config win = b->cfg;
        float win_ms = 1e30f;
        // mean launch time of configuration c over reps launches after a warm-up
        auto time_config = [&](const hipdvr_batch_config& c) {
            b->cfg = c;
            launch_batch(b, false);
            hip_check(hipEventRecord(e0, st), "hipEventRecord");
            for (uint32_t r = 0; r < reps; ++r) launch_batch(b, false);
            hip_check(hipEventRecord(e1, st), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            return ms / (float)reps;
        };
        try {
            for (int t = 0; t < kTileShapes; ++t)
                for (int bl = 3; bl <= 5; ++bl)
                    for (int k : {2, 4}) {
                        const float ms = time_config(hipdvr_batch_config{t, bl, k, 0});
                        if (ms < win_ms) {
                            win_ms = ms;
                            win = b->cfg;
                        }
                    }
            const hipdvr_batch_config shape = win;
            for (int r : {3, 4, 5}) {
                const float ms = time_config(hipdvr_batch_config{shape.tile, shape.brick_log2, shape.batch, r});
                if (ms < win_ms) {
                    win_ms = ms;
                    win = b->cfg;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->cfg = win;
        if (best) *best = win;
        if (best_ms) *best_ms = win_ms;
    });
}

int hipdvr_batch_set_learned(hipdvr_batch* b, hipdvr_learned* const* learned, uint32_t n,
                             const hipdvr_learned_opts* o) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        if (!learned || n != b->jobs.size()) fail(HIPDVR_ERR_INVALID, "one learned state (or NULL) per job");
        ctx_set_device(b->ctx);
        hipStream_t st = ctx_stream(b->ctx);
        const hipdvr_learned_opts lopts = learned_options(o);
        const size_t nj = b->jobs.size();
        // built aside and taken over only when every job is accepted
        std::vector<hipdvr_learned*> lv(nj, nullptr);
        std::vector<char> plain_lv(nj, 0);
        std::vector<const hipdvr_learned*> given(nj, nullptr);   // the state given for each job
        std::vector<JudgeGrid> grids(nj, JudgeGrid{});
        std::vector<float4*> sparse(nj, nullptr);
        std::vector<float*> sse(nj, nullptr);
        std::vector<int*> flag(nj, nullptr);
        bool any = false;
        std::vector<LearnedArgs> largs[kBrickSizes];
        for (int bs = 0; bs < kBrickSizes; ++bs) largs[bs] = b->h_args[bs];
        for (size_t i = 0; i < nj; ++i) {
            hipdvr_learned* l = learned[b->order[i]];
            const hipdvr_batch_job& j = b->jobs[i];
            if (!l) continue;
            const std::string job = "job " + std::to_string(b->order[i]);
            if (learned_volume(l) != j.volume) fail(HIPDVR_ERR_INVALID, job + ": the learned state of another volume");
            if (learned_prepared_tf(l) != j.tf)
                fail(HIPDVR_ERR_INVALID, job + ": the learned state was not prepared for the job's TF");
            if (b->shade == 2) fail(HIPDVR_ERR_UNSUPPORTED, "learned jobs shade with the approximate Phong path");
            // a learned state holds the opacity tables of one step
            for (size_t k = 0; k < i; ++k)
                if (given[k] == l && b->jobs[k].params.step != j.params.step)
                    fail(HIPDVR_ERR_INVALID, "jobs sharing a learned state must share the step");
            given[i] = l;
            if (!learned_gains(l)) {
                // every occupied brick at the base step: the plain march
                // gives the exact frame at a lower cost
                plain_lv[i] = 1;
                continue;
            }
            for (int bs = 0; bs < kBrickSizes; ++bs) learned_stepping_args(l, &j.params, largs[bs][i]);
            lv[i] = l;
            grids[i] = learned_judge_grid(j.w, j.h, lopts);
            any = true;
        }
        for (size_t i = 0; i < nj; ++i) {
            if (!lv[i]) continue;
            if (i < b->d_sparse.size() && b->d_sparse[i] && b->grids[i].w == grids[i].w &&
                b->grids[i].h == grids[i].h) {   // same grid as before: keep its buffers
                sparse[i] = b->d_sparse[i];
                sse[i] = b->d_sse[i];
                flag[i] = b->d_flag[i];
                continue;
            }
            const size_t px = (size_t)grids[i].w * grids[i].h;
            sparse[i] = dev_alloc<float4>(b, px * sizeof(float4));
            sse[i] = dev_alloc<float>(b, 64);
            flag[i] = dev_alloc<int>(b, 64);
        }
        // the learned launch's jobs: each learned view's judge re-march
        // (plain, into its sparse image), then every job; and the judges
        std::vector<LearnedArgs> ljobs[kBrickSizes];
        std::vector<JudgeJob> judges;
        const float mse_max = powf(10.f, -lopts.psnr_floor / 10.f);
        for (size_t i = 0; i < nj; ++i) {
            if (!lv[i]) continue;
            const JudgeGrid& g = grids[i];
            const hipdvr_batch_job& j = b->jobs[i];
            for (int bs = 0; bs < kBrickSizes; ++bs) {
                LearnedArgs sj;
                std::memset(&sj, 0, sizeof(sj));
                sj.m = b->h_args[bs][i].m;
                sj.m.out = sparse[i];
                sj.m.counters = nullptr;
                sj.m.cond_flag = nullptr;
                sj.m.grid_w = g.w;
                sj.m.grid_h = g.h;
                sj.m.stride = (int32_t)g.stride;
                sj.m.pix_block = (int32_t)g.block;
                sj.m.pix_offset = (int32_t)g.offs;
                ljobs[bs].push_back(sj);
            }
            JudgeJob jj;
            jj.learned = (const float4*)j.d_out;
            jj.sparse = sparse[i];
            jj.sse = sse[i];
            jj.flag = flag[i];
            jj.w = (int32_t)j.w;
            jj.h = (int32_t)j.h;
            jj.sw = g.w;
            jj.sh = g.h;
            jj.stride = (int32_t)g.stride;
            jj.block = (int32_t)g.block;
            jj.offs = (int32_t)g.offs;
            jj.denom = (float)(g.w * g.h * 4);
            jj.mse_max = mse_max;
            judges.push_back(jj);
        }
        for (int bs = 0; bs < kBrickSizes; ++bs) ljobs[bs].insert(ljobs[bs].end(), largs[bs].begin(), largs[bs].end());
        const uint32_t nl = (uint32_t)ljobs[0].size();
        std::vector<uint32_t> lstart[kTileShapes];
        for (int t = 0; t < kTileShapes; ++t) {
            const uint32_t tw = 2 * wave_w(t), th = 2 * wave_h(t);
            lstart[t].assign(nl + 1, 0);
            for (uint32_t k = 0; k < nl; ++k) {
                const MarchArgs& m = ljobs[0][k].m;
                const uint64_t tiles = (uint64_t)((m.grid_w + tw - 1) / tw) * ((m.grid_h + th - 1) / th);
                if (lstart[t][k] + tiles > 0xfffffff0ull) fail(HIPDVR_ERR_UNSUPPORTED, "batch too large");
                lstart[t][k + 1] = lstart[t][k] + (uint32_t)tiles;
            }
        }
        // capacity: every job of the batch twice
        for (int bs = 0; bs < kBrickSizes; ++bs)
            if (!b->d_largs[bs]) b->d_largs[bs] = dev_alloc<LearnedArgs>(b, 2 * nj * sizeof(LearnedArgs));
        for (int t = 0; t < kTileShapes; ++t)
            if (!b->d_ltile_start[t]) b->d_ltile_start[t] = dev_alloc<uint32_t>(b, (2 * nj + 1) * 4);
        if (!b->d_judge) b->d_judge = dev_alloc<JudgeJob>(b, nj * sizeof(JudgeJob));
        for (int bs = 0; bs < kBrickSizes; ++bs)
            hip_check(hipMemcpyAsync(b->d_largs[bs], ljobs[bs].data(), nl * sizeof(LearnedArgs), hipMemcpyHostToDevice,
                                     st),
                      "upload");
        for (int t = 0; t < kTileShapes; ++t)
            hip_check(hipMemcpyAsync(b->d_ltile_start[t], lstart[t].data(), (nl + 1) * 4, hipMemcpyHostToDevice, st),
                      "upload");
        if (!judges.empty())
            hip_check(hipMemcpyAsync(b->d_judge, judges.data(), judges.size() * sizeof(JudgeJob), hipMemcpyHostToDevice,
                                     st),
                      "upload");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        b->n_ljobs = nl;
        for (int t = 0; t < kTileShapes; ++t) b->ltotal_tiles[t] = lstart[t][nl];
        b->n_judge = (uint32_t)judges.size();
        b->lopts = lopts;
        b->learned.swap(lv);
        b->plain_learned.swap(plain_lv);
        b->grids.swap(grids);
        b->d_sparse.swap(sparse);
        b->d_sse.swap(sse);
        b->d_flag.swap(flag);
    });
}

int hipdvr_batch_launch_learned(hipdvr_batch* b, hipdvr_learned_stats* stats) {
    if (!b) return HIPDVR_ERR_INVALID;
    return ctx_guarded(b->ctx, [&] {
        ctx_set_device(b->ctx);
        hipStream_t st = ctx_stream(b->ctx);
        const size_t nj = b->jobs.size();
        const int bs = b->cfg.brick_log2 - 3;
        const int shade = b->shade;
        bool any = false;
        for (size_t i = 0; i < nj; ++i) {
            if (!b->learned[i]) continue;
            any = true;
            // the state's opacity tables follow the last step it rendered with
            learned_ensure_step(b->learned[i], b->jobs[i].params.step);
        }
        const bool count = stats != nullptr;
        if (count) launch_zero_u64(b->ctx->d_counters, kNumCounters, st);
        const bool timed = ctx_timing_start(b->ctx);
        if (!b->d_largs[0]) fail(HIPDVR_ERR_INVALID, "hipdvr_batch_set_learned has not been called");
        // 1. the judges' plain re-marches of the learned views' sparse pixels
        //    and every job, learned or plain, in one persistent launch
        launch_batch(b, count, true);
        if (any) {
            // 2. the judges of all learned views in one launch, then per
            //    learned view the plain frame, launched always, which runs
            //    only when its view was flagged
            launch_judge_multi(b->d_judge, b->n_judge, st);
            for (size_t i = 0; i < nj; ++i) {
                if (!b->learned[i]) continue;
                const hipdvr_batch_job& j = b->jobs[i];
                MarchArgs fb = b->h_args[bs][i].m;
                fb.counters = nullptr;
                fb.cond_flag = b->d_flag[i];
                launch_march_cell(fb, fb.dtype, shade, false, j.volume->cells_wide, default_batch(shade), st);
            }
        }
        hip_check(hipGetLastError(), "batched learned launch");
        if (timed) ctx_timing_stop(b->ctx);
        if (count) {
            unsigned long long c[kNumCounters];
            std::vector<float> sse(nj, 0.f);
            std::vector<int> flag(nj, 0);
            hip_check(hipMemcpyAsync(c, b->ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st), "download");
            for (size_t i = 0; i < nj; ++i) {
                if (!b->learned[i]) continue;
                hip_check(hipMemcpyAsync(&sse[i], b->d_sse[i], 4, hipMemcpyDeviceToHost, st), "download");
                hip_check(hipMemcpyAsync(&flag[i], b->d_flag[i], 4, hipMemcpyDeviceToHost, st), "download");
            }
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            for (size_t i = 0; i < nj; ++i) {
                hipdvr_learned_stats& s = stats[b->order[i]];
                s.samples = c[1];
                s.shaded = c[3];
                if (b->learned[i]) {
                    const JudgeGrid& g = b->grids[i];
                    const double mse = (double)sse[i] / ((double)g.w * g.h * 4);
                    s.sparse_psnr = mse > 0 ? (float)(10.0 * std::log10(1.0 / mse)) : 999.f;
                    s.fell_back = flag[i];
                } else {
                    s.sparse_psnr = b->plain_learned[i] ? 999.f : 0.f;
                    s.fell_back = 0;
                }
            }
        }
    });
}

}  // extern "C"
