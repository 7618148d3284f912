// march_cell.hip -- per-frame launches of the cell-record ray march (the
// device code and its parity argument are in march_cell.h): one workgroup per
// 16x16 pixel tile, tiles in XCD-compact order.
#include "kernels.h"
#include "march_cell.h"
#include "march_common.h"

namespace hipdvr {

template <typename T, bool WIDE, int SHADE, bool COUNT, int K>
__global__ __launch_bounds__(256) void march_cell_kernel(const MarchArgs a, const uint32_t tiles_x,
                                                         const uint32_t tiles_y) {
    // XCD-compact tile order (xcd_tile): each XCD marches one compact region
    // of tiles, the chip one compact image area
    uint32_t tile_x, tile_y;
    if (!xcd_tile(blockIdx.x, tiles_x, tiles_y, tile_x, tile_y)) return;   // whole workgroup leaves together
    if (a.cond_flag && *a.cond_flag == 0) return;
    __shared__ uint32_t s_tf[kMaxTfWidth];
    __shared__ float s_alpha[256];
    extern __shared__ uint32_t s_occ[];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.tf_width; i += 256) s_tf[i] = a.tf[i];
    s_alpha[tid] = a.alpha_corr[tid];
    const bool occ_lds = a.occ_words <= kMaxLdsOccWords;
    if (occ_lds)
        for (uint32_t i = tid; i < a.occ_words; i += 256) s_occ[i] = a.occ_bits[i];
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int gx = (int)tile_x * 16 + (wave & 1) * 8 + (lane & 7);
    const int gy = (int)tile_y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int px = gx * a.stride + a.pix_offset, py = gy * a.stride + a.pix_offset;
    const bool inside = (gx < a.grid_w) && (gy < a.grid_h) && (px < a.w) && (py < a.h);
    const CellFetch<T, WIDE> f = cell_fetch_of<T, WIDE>(a);
    RayCounters n;
    if (inside) {
        const size_t pix = (size_t)py * a.w + px;
        const float4 e = a.entry[pix];
        const float4 l = a.exit[pix];
        float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!is_background(e, l)) res = cell_ray<T, WIDE, SHADE, COUNT, K>(a, f, s_tf, s_alpha, occ_lds ? s_occ : nullptr,
                                                                          px, py, e, l, n);
        a.out[(size_t)gy * a.grid_w + gx] = res;
    }
    if (COUNT) {
        wave_add(a.counters + 0, n.rays);
        wave_add(a.counters + 1, n.samples);
        wave_add(a.counters + 2, n.fetched);
        wave_add(a.counters + 3, n.shaded);
        wave_add_max64(a.counters + 4, n.slots);
        wave_add_max64(a.counters + 5, n.shade_slots);
    }
}

template <typename T, bool WIDE, int SHADE, int K>
static void launch_k(const MarchArgs& a, bool count, hipStream_t st) {
    const uint32_t tx = (a.grid_w + 15) / 16, ty = (a.grid_h + 15) / 16;
    const dim3 grid(xcd_tile_grid(tx, ty));
    size_t shm = a.occ_words <= kMaxLdsOccWords ? (size_t)a.occ_words * 4 : 0;
    if (SHADE == 0) shm = unshaded_dyn_lds(shm);   // five workgroups per CU
    if (count)
        hipLaunchKernelGGL((march_cell_kernel<T, WIDE, SHADE, true, K>), grid, dim3(256), shm, st, a, tx, ty);
    else
        hipLaunchKernelGGL((march_cell_kernel<T, WIDE, SHADE, false, K>), grid, dim3(256), shm, st, a, tx, ty);
}

template <typename T, bool WIDE, int SHADE>
static void launch_ks(const MarchArgs& a, bool count, int batch, hipStream_t st) {
    if (batch == 1) launch_k<T, WIDE, SHADE, 1>(a, count, st);
    else if (batch == 2) launch_k<T, WIDE, SHADE, 2>(a, count, st);
    else launch_k<T, WIDE, SHADE, 4>(a, count, st);
}

template <typename T, bool WIDE>
static void launch_cw(const MarchArgs& a, int shade, bool count, int batch, hipStream_t st) {
    if (shade == 2) launch_ks<T, WIDE, 2>(a, count, batch, st);
    else if (shade == 1) launch_ks<T, WIDE, 1>(a, count, batch, st);
    else launch_ks<T, WIDE, 0>(a, count, batch, st);
}

void launch_march_cell(const MarchArgs& a, int dtype, int shade, bool count, bool wide, int batch,
                       hipStream_t st) {
    if (dtype == 0) {
        if (wide) launch_cw<uint8_t, true>(a, shade, count, batch, st);
        else launch_cw<uint8_t, false>(a, shade, count, batch, st);
    } else {
        if (wide) launch_cw<uint16_t, true>(a, shade, count, batch, st);
        else launch_cw<uint16_t, false>(a, shade, count, batch, st);
    }
}

}  // namespace hipdvr
