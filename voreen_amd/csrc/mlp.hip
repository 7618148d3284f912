// mlp.hip -- the per-brick stepping network: 16 -> 64 -> 64 -> 2, GELU,
// BF16 weights, FP32 accumulation on MFMA (v_mfma_f32_16x16x16_bf16).
//
// One wavefront handles tiles of 16 rows.  For a 16x16x16 MFMA the operand
// layout is: A[m][k] lives in lane m + 16*(k/4), element k%4; B[k][n] in lane
// n + 16*(k/4), element k%4; the accumulator D[m][n] in lane n + 16*(m/4),
// element m%4.  Activations go through LDS between layers to move from the
// accumulator layout to the A layout.
//
// Parameters (FP32 master copy, flat): W1[16][64] b1[64] W2[64][64] b2[64]
// W3[64][2] b3[2]; the kernels read BF16 copies of W1, W2, W3 (W3 padded to
// 64x16 with zero columns) made by mlp_pack after every update.
//
// Loss per row (targets t0 in {0,1}, t1 in [0,3], weight t2):
//   BCE-with-logits(y0, t0) + t2 * (y1 - t1)^2, averaged over the batch.
#include <hip/hip_runtime.h>

#include "mlp.h"

namespace hipdvr {

typedef __attribute__((ext_vector_type(4))) float f4;
typedef __attribute__((ext_vector_type(4))) short s4;

__device__ __forceinline__ short f2bf(float x) {
    const uint32_t u = __float_as_uint(x);
    const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;   // round to nearest even
    return (short)((u + r) >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// GELU (erf form) and its derivative.
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Shared weights of one workgroup (BF16) and biases.
struct MlpShared {
    short w1[kMlpIn * kMlpHidden];        // [16][64]
    short w2[kMlpHidden * kMlpHidden];    // [64][64]
    short w3[kMlpHidden * 16];            // [64][16] (columns >= 2 are zero)
    float b1[kMlpHidden], b2[kMlpHidden], b3[16];
};

// Per-wave scratch: activations of the 16-row tile.
struct WaveTile {
    short x[16 * kMlpIn];       // input rows (bf16)
    short h1[16 * kMlpHidden];  // GELU(z1) (bf16)
    short h2[16 * kMlpHidden];  // GELU(z2) (bf16)
    float z1[16 * kMlpHidden];
    float z2[16 * kMlpHidden];
    float dz[16 * kMlpHidden];  // dZ2, then dZ1
    float dy[16 * 2];
};

__device__ void load_shared(MlpShared& sh, const MlpDevParams& p) {
    for (int i = threadIdx.x; i < kMlpIn * kMlpHidden; i += blockDim.x) sh.w1[i] = p.w1[i];
    for (int i = threadIdx.x; i < kMlpHidden * kMlpHidden; i += blockDim.x) sh.w2[i] = p.w2[i];
    for (int i = threadIdx.x; i < kMlpHidden * 16; i += blockDim.x) sh.w3[i] = p.w3[i];
    for (int i = threadIdx.x; i < kMlpHidden; i += blockDim.x) {
        sh.b1[i] = p.b[i];
        sh.b2[i] = p.b[kMlpHidden + i];
    }
    for (int i = threadIdx.x; i < 16; i += blockDim.x) sh.b3[i] = i < kMlpOut ? p.b[2 * kMlpHidden + i] : 0.f;
}

// Forward pass of one 16-row tile whose inputs are already in t.x.
// Returns y0/y1 of the lane's rows in y (lanes with n = lane%16 < 2 hold
// output column n for rows 4*(lane/16) .. +3).
__device__ void tile_forward(const MlpShared& sh, WaveTile& t, f4& y) {
    const int lane = threadIdx.x & 63;
    const int n0 = lane & 15, kq = (lane >> 4) * 4;
    // layer 1: [16x16] x [16x64]
    s4 a;
    for (int i = 0; i < 4; ++i) a[i] = t.x[n0 * kMlpIn + kq + i];
    for (int nt = 0; nt < 4; ++nt) {
        s4 b;
        for (int i = 0; i < 4; ++i) b[i] = sh.w1[(kq + i) * kMlpHidden + nt * 16 + n0];
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(a, b, acc);
        for (int i = 0; i < 4; ++i) {
            const int row = kq + i, col = nt * 16 + n0;
            const float z = acc[i] + sh.b1[col];
            t.z1[row * kMlpHidden + col] = z;
            t.h1[row * kMlpHidden + col] = f2bf(gelu(z));
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 2: [16x64] x [64x64]
    for (int nt = 0; nt < 4; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < 4; ++kc) {
            s4 aa, b;
            for (int i = 0; i < 4; ++i) {
                aa[i] = t.h1[n0 * kMlpHidden + kc * 16 + kq + i];
                b[i] = sh.w2[(kc * 16 + kq + i) * kMlpHidden + nt * 16 + n0];
            }
            acc = mfma(aa, b, acc);
        }
        for (int i = 0; i < 4; ++i) {
            const int row = kq + i, col = nt * 16 + n0;
            const float z = acc[i] + sh.b2[col];
            t.z2[row * kMlpHidden + col] = z;
            t.h2[row * kMlpHidden + col] = f2bf(gelu(z));
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 3: [16x64] x [64x16]
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < 4; ++kc) {
        s4 aa, b;
        for (int i = 0; i < 4; ++i) {
            aa[i] = t.h2[n0 * kMlpHidden + kc * 16 + kq + i];
            b[i] = sh.w3[(kc * 16 + kq + i) * 16 + n0];
        }
        acc = mfma(aa, b, acc);
    }
    for (int i = 0; i < 4; ++i) y[i] = acc[i] + sh.b3[n0];
}

// Load 16 input rows (zero rows past `rows`) into the tile as BF16.
__device__ void tile_load_x(WaveTile& t, const float* x, uint32_t row0, uint32_t rows) {
    const int lane = threadIdx.x & 63;
    for (int e = lane; e < 16 * kMlpIn; e += 64) {
        const uint32_t r = row0 + e / kMlpIn;
        t.x[e] = f2bf(r < rows ? x[(size_t)r * kMlpIn + e % kMlpIn] : 0.f);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpDevParams p, const float* __restrict__ x,
                                                          uint32_t rows, float* __restrict__ y) {
    __shared__ MlpShared sh;
    __shared__ WaveTile tiles[4];
    load_shared(sh, p);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveTile& t = tiles[wave];
    const uint32_t ntiles = (rows + 15) / 16;
    for (uint32_t tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        tile_load_x(t, x, tile * 16, rows);
        f4 out;
        tile_forward(sh, t, out);
        const int n0 = lane & 15, kq = (lane >> 4) * 4;
        if (n0 < kMlpOut)
            for (int i = 0; i < 4; ++i) {
                const uint32_t r = tile * 16 + kq + i;
                if (r < rows) y[(size_t)r * kMlpOut + n0] = out[i];
            }
        __builtin_amdgcn_wave_barrier();
    }
}

// Forward + backward over the batch; accumulates dL/dparam (FP32) into grad
// and the summed loss into loss[0].  Each wavefront keeps its weight-gradient
// partial sums in accumulator registers over its tiles and adds them to
// global memory once.
__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpDevParams p, const float* __restrict__ x,
                                                       const float* __restrict__ tgt, uint32_t rows,
                                                       float inv_batch, float* __restrict__ grad,
                                                       float* __restrict__ loss) {
    __shared__ MlpShared sh;
    __shared__ WaveTile tiles[4];
    load_shared(sh, p);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n0 = lane & 15, kq = (lane >> 4) * 4;
    WaveTile& t = tiles[wave];
    f4 gw2[4][4], gw1[4];
    for (int a = 0; a < 4; ++a) {
        gw1[a] = f4{0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < 4; ++b) gw2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    }
    float gw3[2] = {0.f, 0.f};     // lane = hidden unit k: dW3[k][0..1]
    float gb1 = 0.f, gb2 = 0.f;    // lane = hidden unit
    float gb3[2] = {0.f, 0.f};
    float lsum = 0.f;
    const uint32_t ntiles = (rows + 15) / 16;
    for (uint32_t tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        tile_load_x(t, x, tile * 16, rows);
        f4 y;
        tile_forward(sh, t, y);
        // dL/dy
        if (n0 < kMlpOut)
            for (int i = 0; i < 4; ++i) {
                const uint32_t r = tile * 16 + kq + i;
                float d = 0.f;
                if (r < rows) {
                    const float* tr = tgt + (size_t)r * kMlpTargets;
                    if (n0 == 0) {
                        const float s = 1.f / (1.f + expf(-y[i]));
                        d = s - tr[0];
                        lsum += fmaxf(y[i], 0.f) - y[i] * tr[0] + log1pf(expf(-fabsf(y[i])));
                    } else {
                        const float e = y[i] - tr[1];
                        d = 2.f * tr[2] * e;
                        lsum += tr[2] * e * e;
                    }
                }
                t.dy[(kq + i) * 2 + n0] = d * inv_batch;
            }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // layer 3 gradients (K = 2: scalar); lane = hidden unit k
        {
            const int k = lane;
            const float w30 = bf2f(sh.w3[k * 16 + 0]), w31 = bf2f(sh.w3[k * 16 + 1]);
            for (int r = 0; r < 16; ++r) {
                const float d0 = t.dy[r * 2], d1 = t.dy[r * 2 + 1];
                const float h = bf2f(t.h2[r * kMlpHidden + k]);
                gw3[0] += h * d0;
                gw3[1] += h * d1;
                const float dh = d0 * w30 + d1 * w31;
                const float dz = dh * gelu_grad(t.z2[r * kMlpHidden + k]);
                t.dz[r * kMlpHidden + k] = dz;
                gb2 += dz;
            }
            if (lane < 2)
                for (int r = 0; r < 16; ++r) gb3[lane] += t.dy[r * 2 + lane];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW2[k][j] += sum_r h1[r][k] dz2[r][j]   (M = k, N = j, K = r)
        for (int mt = 0; mt < 4; ++mt) {
            s4 a;
            for (int i = 0; i < 4; ++i) a[i] = t.h1[(kq + i) * kMlpHidden + mt * 16 + n0];
            for (int nt = 0; nt < 4; ++nt) {
                s4 b;
                for (int i = 0; i < 4; ++i) b[i] = f2bf(t.dz[(kq + i) * kMlpHidden + nt * 16 + n0]);
                gw2[mt][nt] = mfma(a, b, gw2[mt][nt]);
            }
        }
        // dH1[r][k] = sum_j dz2[r][j] W2[k][j]   (M = r, N = k, K = j)
        f4 dh1[4];
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kc = 0; kc < 4; ++kc) {
                s4 a, b;
                for (int i = 0; i < 4; ++i) {
                    a[i] = f2bf(t.dz[n0 * kMlpHidden + kc * 16 + kq + i]);
                    b[i] = sh.w2[(nt * 16 + n0) * kMlpHidden + kc * 16 + kq + i];
                }
                acc = mfma(a, b, acc);
            }
            dh1[nt] = acc;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dZ1 = dH1 * gelu'(z1)
        for (int nt = 0; nt < 4; ++nt)
            for (int i = 0; i < 4; ++i) {
                const int row = kq + i, col = nt * 16 + n0;
                t.dz[row * kMlpHidden + col] = dh1[nt][i] * gelu_grad(t.z1[row * kMlpHidden + col]);
            }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int r = 0; r < 16; ++r) gb1 += t.dz[r * kMlpHidden + lane];
        // dW1[f][k] += sum_r x[r][f] dz1[r][k]   (M = f, N = k, K = r)
        {
            s4 a;
            for (int i = 0; i < 4; ++i) a[i] = t.x[(kq + i) * kMlpIn + n0];
            for (int nt = 0; nt < 4; ++nt) {
                s4 b;
                for (int i = 0; i < 4; ++i) b[i] = f2bf(t.dz[(kq + i) * kMlpHidden + nt * 16 + n0]);
                gw1[nt] = mfma(a, b, gw1[nt]);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
    // flush this wavefront's partial sums
    float* gW1 = grad;
    float* gB1 = grad + kMlpIn * kMlpHidden;
    float* gW2 = gB1 + kMlpHidden;
    float* gB2 = gW2 + kMlpHidden * kMlpHidden;
    float* gW3 = gB2 + kMlpHidden;
    float* gB3 = gW3 + kMlpHidden * kMlpOut;
    for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i)
            atomicAdd(&gW1[(kq + i) * kMlpHidden + nt * 16 + n0], gw1[nt][i]);
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt)
            for (int i = 0; i < 4; ++i)
                atomicAdd(&gW2[(mt * 16 + kq + i) * kMlpHidden + nt * 16 + n0], gw2[mt][nt][i]);
    atomicAdd(&gB1[lane], gb1);
    atomicAdd(&gB2[lane], gb2);
    atomicAdd(&gW3[lane * kMlpOut + 0], gw3[0]);
    atomicAdd(&gW3[lane * kMlpOut + 1], gw3[1]);
    if (lane < 2) atomicAdd(&gB3[lane], gb3[lane]);
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) atomicAdd(loss, lsum * inv_batch);
}

// Adam step on the FP32 master parameters.
__global__ void mlp_adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ g, int n, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master -> BF16 weights (W3 padded to 16 columns) + FP32 biases.
__global__ void mlp_pack_kernel(const float* __restrict__ w, MlpDevParams p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = w;
    const float* B1 = W1 + kMlpIn * kMlpHidden;
    const float* W2 = B1 + kMlpHidden;
    const float* B2 = W2 + kMlpHidden * kMlpHidden;
    const float* W3 = B2 + kMlpHidden;
    const float* B3 = W3 + kMlpHidden * kMlpOut;
    if (i < kMlpIn * kMlpHidden) p.w1[i] = f2bf(W1[i]);
    if (i < kMlpHidden * kMlpHidden) p.w2[i] = f2bf(W2[i]);
    if (i < kMlpHidden * 16) {
        const int k = i / 16, j = i % 16;
        p.w3[i] = j < kMlpOut ? f2bf(W3[k * kMlpOut + j]) : (short)0;
    }
    if (i < kMlpHidden) {
        p.b[i] = B1[i];
        p.b[kMlpHidden + i] = B2[i];
    }
    if (i < kMlpOut) p.b[2 * kMlpHidden + i] = B3[i];
}

__global__ void zero_f32_kernel(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

void launch_mlp_pack(const float* master, const MlpDevParams& p, hipStream_t st) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kMlpHidden * kMlpHidden + 255) / 256), dim3(256), 0, st,
                       master, p);
}

void launch_mlp_forward(const MlpDevParams& p, const float* x, uint32_t rows, float* y, hipStream_t st) {
    const uint32_t ntiles = (rows + 15) / 16;
    uint32_t blocks = (ntiles + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    if (blocks == 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, st, p, x, rows, y);
}

void launch_mlp_grad(const MlpDevParams& p, const float* x, const float* tgt, uint32_t rows,
                     float* grad, float* loss, hipStream_t st) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st, grad,
                       kMlpParams);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, st, loss, 1);
    const uint32_t ntiles = (rows + 15) / 16;
    uint32_t blocks = (ntiles + 3) / 4;
    if (blocks > 256) blocks = 256;
    if (blocks == 0) return;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, st, p, x, tgt, rows,
                       1.0f / (float)rows, grad, loss);
}

void launch_mlp_adam(float* w, float* m, float* v, const float* g, float lr, int step, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, m, v, g,
                       kMlpParams, lr, b1, b2, 1e-8f, c1, c2);
}

}  // namespace hipdvr
