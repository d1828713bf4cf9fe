// yz_mlp.hip -- the learned preconditioner's network and its training,
// on the MFMA units of gfx950 (see yz_mlp.h for the model).
//
// MFMA layout of v_mfma_f32_16x16x16_bf16 (one wave, 64 lanes, lane l):
//   A (16 x 16, rows m, depth k):  a[j] = A[l % 16][4 (l / 16) + j]
//   B (16 x 16, depth k, cols n):  b[j] = B[4 (l / 16) + j][l % 16]
//   C (16 x 16):                   c[j] = C[4 (l / 16) + j][l % 16]
// One wave takes 16 cells at a time. Training: the 16 x 8 features (padded
// to 16) times W1^T, then 16 x 64 times W2^T and W3^T, the activations
// passing through the wave's own LDS tile between layers. (Inference runs
// the product transposed, with no LDS, k_mlp_forward.) The backward pass
// recomputes the forward for its tile, forms dY from dL/dM, and accumulates
// the weight gradients over all its tiles in MFMA accumulators; each wave
// writes its partial gradients once, and one kernel sums them in a fixed
// order. Each wave works on LDS tiles of its own, so waves synchronise
// among their own lanes only (the block barrier is for the weights).
#include "yz_mlp.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace yz {

#define MLP_CHECK(x)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int MBLK = 256;          // 4 waves
static constexpr int WAVES = MBLK / 64;
static constexpr int LDT = 68;            // padded row of a 16 x 64 activation tile (floats)
static constexpr int LDY = 20;            // padded row of the 16 x 16 output tile
static constexpr int TRAIN_BLOCKS = 512;  // blocks of the backward pass (2048 waves of partials)

// round to nearest even (finite inputs)
__device__ inline short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// the wave's own LDS tile written by some lanes and read by others: order
// the accesses within the wave (each wave has private tiles, so no block
// barrier is needed)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WLds {
    short w1[64 * 16];
    short w2[64 * 64];
    short w3[16 * 64];
    float b1[64], b2[64], b3[16];
};

__device__ inline void load_weights(const float *__restrict__ P, WLds &W) {
    for (int t = threadIdx.x; t < 64 * 16; t += blockDim.x) W.w1[t] = f2bf(P[MLP_W1 + t]);
    for (int t = threadIdx.x; t < 64 * 64; t += blockDim.x) W.w2[t] = f2bf(P[MLP_W2 + t]);
    for (int t = threadIdx.x; t < 16 * 64; t += blockDim.x) W.w3[t] = f2bf(P[MLP_W3 + t]);
    for (int t = threadIdx.x; t < 64; t += blockDim.x) { W.b1[t] = P[MLP_B1 + t]; W.b2[t] = P[MLP_B2 + t]; }
    for (int t = threadIdx.x; t < 16; t += blockDim.x) W.b3[t] = P[MLP_B3 + t];
}

// A fragment of k-block kb from a tile T[16][ld] (rows = cells)
__device__ inline bf16x4 fragA(const float *T, int ld, int kb, int l) {
    bf16x4 a;
    const float *r = T + (l & 15) * ld + 16 * kb + 4 * (l >> 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = f2bf(r[j]);
    return a;
}
// fragment of block cb of a tile read "by columns": element j is
// T[4 (l/16) + j][16 cb + l % 16] -- the B operand B[k=m][n] = T[m][n], or
// the A operand of the transpose, A[o][k=m] = T[m][o]
__device__ inline bf16x4 fragT(const float *T, int ld, int cb, int l) {
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = f2bf(T[(4 * (l >> 4) + j) * ld + 16 * cb + (l & 15)]);
    return a;
}
// B fragment for x W^T: B[k=i][n=o] = W[o][i], W row-major [O][I]
__device__ inline bf16x4 fragWf(const short *W, int I, int nb, int kb, int l) {
    bf16x4 b;
    const short *r = W + (16 * nb + (l & 15)) * I + 16 * kb + 4 * (l >> 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = r[j];
    return b;
}
// B fragment for d W: B[k=o][n=i] = W[o][i]
__device__ inline bf16x4 fragWb(const short *W, int I, int ib, int kb, int l) {
    bf16x4 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = W[(16 * kb + 4 * (l >> 4) + j) * I + 16 * ib + (l & 15)];
    return b;
}
// features of cells c0..c0+15 as the A operand (8 features, padded to 16)
__device__ inline bf16x4 fragX(const float *__restrict__ feat, int N, int c0, int l) {
    bf16x4 a;
    const int cell = c0 + (l & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = 4 * (l >> 4) + j;
        a[j] = f2bf((f < MLP_IN && cell < N) ? feat[(size_t)f * N + cell] : 0.0f);
    }
    return a;
}
// the same features as the B operand, B[k=m][n=f] = X[m][f]
__device__ inline bf16x4 fragXB(const float *__restrict__ feat, int N, int c0, int l) {
    bf16x4 b;
    const int f = l & 15;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int cell = c0 + 4 * (l >> 4) + j;
        b[j] = f2bf((f < MLP_IN && cell < N) ? feat[(size_t)f * N + cell] : 0.0f);
    }
    return b;
}

// forward of one 16-cell tile: H1, H2 into the wave's tiles, Y returned
// in C layout (Y[4 (l/16) + j][l % 16])
__device__ inline f32x4 tile_forward(const WLds &W, const float *__restrict__ feat, int N, int c0, float *H1,
                                     float *H2, int l) {
    const bf16x4 ax = fragX(feat, N, c0, l);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        f32x4 acc = {0, 0, 0, 0};
        acc = mfma(ax, fragWf(W.w1, 16, nb, 0, l), acc);
        const float bias = W.b1[16 * nb + (l & 15)];
#pragma unroll
        for (int j = 0; j < 4; ++j) H1[(4 * (l >> 4) + j) * LDT + 16 * nb + (l & 15)] = fmaxf(acc[j] + bias, 0.0f);
    }
    wave_sync();
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(fragA(H1, LDT, kb, l), fragWf(W.w2, 64, nb, kb, l), acc);
        const float bias = W.b2[16 * nb + (l & 15)];
#pragma unroll
        for (int j = 0; j < 4; ++j) H2[(4 * (l >> 4) + j) * LDT + 16 * nb + (l & 15)] = fmaxf(acc[j] + bias, 0.0f);
    }
    wave_sync();
    f32x4 y = {0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) y = mfma(fragA(H2, LDT, kb, l), fragWf(W.w3, 64, 0, kb, l), y);
    const float b3 = W.b3[l & 15];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] += b3;
    return y;
}

// Inference runs the network transposed, H^T = W X^T: the weights are the
// A operands, held in registers for the whole kernel, and the activations
// of 16 cells are the B operands. The C layout of one layer's output,
// C[4 (l/16) + j][l % 16] (features x cells), is the B layout of the next
// layer's input, so the activations never leave the registers: no LDS tile
// and no lane exchange between the layers.
//
// A fragment of W (row-major [O][I]) for output block ob, input block kb:
// a[j] = W[16 ob + l % 16][16 kb + 4 (l / 16) + j]
__device__ inline bf16x4 fragWa(const short *W, int I, int ob, int kb, int l) {
    bf16x4 a;
    const short *r = W + (16 * ob + (l & 15)) * I + 16 * kb + 4 * (l >> 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = r[j];
    return a;
}
// the features of cells c0..c0+15 as the B operand of X^T:
// b[j] = feat[4 (l / 16) + j][c0 + l % 16] (features 8..15 zero)
__device__ inline bf16x4 fragXT(const float *__restrict__ feat, int N, int c0, int l) {
    bf16x4 b;
    const int cell = c0 + (l & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = 4 * (l >> 4) + j;
        b[j] = f2bf((f < MLP_IN && cell < N) ? feat[(size_t)f * N + cell] : 0.0f);
    }
    return b;
}

// M[o][cell] = Y[cell][o] * mscale[cell], o < 13 (FP32); or, with HALF, the
// raw outputs Y[o][cell] in FP16 (the row scale is applied by the consumer)
template <bool HALF>
__global__ void __launch_bounds__(MBLK) k_mlp_forward(int N, const float *__restrict__ P,
                                                      const float *__restrict__ feat,
                                                      const float *__restrict__ mscale, float *__restrict__ M,
                                                      __half *__restrict__ Yh) {
    __shared__ WLds W;
    load_weights(P, W);
    __syncthreads();
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = 4 * (l >> 4);   // the rows (features, outputs) this lane holds: r0 .. r0 + 3
    bf16x4 w1[4], w2[16], w3[4];
    float b1[4][4], b2[4][4], b3[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        w1[ob] = fragWa(W.w1, 16, ob, 0, l);
        w3[ob] = fragWa(W.w3, 64, 0, ob, l);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) w2[ob * 4 + kb] = fragWa(W.w2, 64, ob, kb, l);
#pragma unroll
        for (int j = 0; j < 4; ++j) { b1[ob][j] = W.b1[16 * ob + r0 + j]; b2[ob][j] = W.b2[16 * ob + r0 + j]; }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) b3[j] = W.b3[r0 + j];
    const int ntiles = (N + 15) / 16;
    const int stride = gridDim.x * WAVES;
    int tile = blockIdx.x * WAVES + wv;
    bf16x4 xb = fragXT(feat, N, tile * 16, l);
    for (; tile < ntiles; tile += stride) {
        const int c0 = tile * 16;
        // the next tile's features, in flight while this tile computes
        const bf16x4 xn = fragXT(feat, N, (tile + stride) * 16, l);
        bf16x4 h1[4], h2[4];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            f32x4 acc = {0, 0, 0, 0};
            acc = mfma(w1[ob], xb, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) h1[ob][j] = f2bf(fmaxf(acc[j] + b1[ob][j], 0.0f));
        }
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma(w2[ob * 4 + kb], h1[kb], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) h2[ob][j] = f2bf(fmaxf(acc[j] + b2[ob][j], 0.0f));
        }
        f32x4 y = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) y = mfma(w3[kb], h2[kb], y);
        const int cell = c0 + (l & 15);
        if (cell < N) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int o = r0 + j;
                if (o < MLP_OUT) {
                    const float v = y[j] + b3[j];
                    if (HALF) Yh[(size_t)o * N + cell] = __float2half(v);
                    else M[(size_t)o * N + cell] = v * mscale[cell];
                }
            }
        }
        xb = xn;
    }
}

// backward: dL/dM (G, 13 x N) -> partial gradients of every parameter,
// one set per wave (partial[wave][MLP_NP])
__global__ void __launch_bounds__(MBLK) k_mlp_backward(int N, const float *__restrict__ P,
                                                       const float *__restrict__ feat,
                                                       const float *__restrict__ mscale,
                                                       const float *__restrict__ G, float *__restrict__ partial) {
    __shared__ WLds W;
    __shared__ float tiles[WAVES][3][16 * LDT];
    __shared__ float ytile[WAVES][16 * LDY];
    load_weights(P, W);
    __syncthreads();  // (the only block barrier: the weights)
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float *H1 = tiles[wv][0], *H2 = tiles[wv][1], *DZ = tiles[wv][2], *DY = ytile[wv];
    f32x4 gW3[4], gW2[16], gW1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { gW3[q] = (f32x4){0, 0, 0, 0}; gW1[q] = (f32x4){0, 0, 0, 0}; }
#pragma unroll
    for (int q = 0; q < 16; ++q) gW2[q] = (f32x4){0, 0, 0, 0};
    float gb1 = 0, gb2 = 0, gb3 = 0;
    const int ntiles = (N + 15) / 16;
    for (int tile = blockIdx.x * WAVES + wv; tile < ntiles; tile += gridDim.x * WAVES) {
        const int c0 = tile * 16;
        (void)tile_forward(W, feat, N, c0, H1, H2, l);
        // dY[m][o] = G[o][cell] mscale[cell] (0 past the 13 outputs or the grid)
        for (int t = l; t < 16 * 16; t += 64) {
            const int m = t >> 4, o = t & 15;
            const int cell = c0 + m;
            DY[m * LDY + o] = (o < MLP_OUT && cell < N) ? G[(size_t)o * N + cell] * mscale[cell] : 0.0f;
        }
        wave_sync();
        // dW3 += dY^T H2, db3 += sum dY
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) gW3[ib] = mfma(fragT(DY, LDY, 0, l), fragT(H2, LDT, ib, l), gW3[ib]);
        if (l < 16)
            for (int m = 0; m < 16; ++m) gb3 += DY[m * LDY + l];
        // dZ2 = (dY W3) * [H2 > 0]
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            f32x4 acc = {0, 0, 0, 0};
            acc = mfma(fragA(DY, LDY, 0, l), fragWb(W.w3, 64, ib, 0, l), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = 4 * (l >> 4) + j, i = 16 * ib + (l & 15);
                DZ[m * LDT + i] = H2[m * LDT + i] > 0.0f ? acc[j] : 0.0f;
            }
        }
        wave_sync();
        // dW2 += dZ2^T H1, db2 += sum dZ2
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
                gW2[ob * 4 + ib] = mfma(fragT(DZ, LDT, ob, l), fragT(H1, LDT, ib, l), gW2[ob * 4 + ib]);
        for (int m = 0; m < 16; ++m) gb2 += DZ[m * LDT + l];
        // dZ1 = (dZ2 W2) * [H1 > 0], into H2's tile (no longer needed)
        f32x4 dz1[4];
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma(fragA(DZ, LDT, kb, l), fragWb(W.w2, 64, ib, kb, l), acc);
            dz1[ib] = acc;
        }
        wave_sync();
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = 4 * (l >> 4) + j, i = 16 * ib + (l & 15);
                H2[m * LDT + i] = H1[m * LDT + i] > 0.0f ? dz1[ib][j] : 0.0f;
            }
        wave_sync();
        // dW1 += dZ1^T X, db1 += sum dZ1
        const bf16x4 xb = fragXB(feat, N, c0, l);
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) gW1[ob] = mfma(fragT(H2, LDT, ob, l), xb, gW1[ob]);
        for (int m = 0; m < 16; ++m) gb1 += H2[m * LDT + l];
        wave_sync();
    }
    // this wave's partial gradients
    float *out = partial + (size_t)(blockIdx.x * WAVES + wv) * MLP_NP;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * (l >> 4) + j, i = 16 * ib + (l & 15);
            out[MLP_W3 + o * 64 + i] = gW3[ib][j];
        }
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int o = 16 * ob + 4 * (l >> 4) + j, i = 16 * ib + (l & 15);
                out[MLP_W2 + o * 64 + i] = gW2[ob * 4 + ib][j];
            }
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 16 * ob + 4 * (l >> 4) + j, f = l & 15;
            out[MLP_W1 + o * 16 + f] = gW1[ob][j];
        }
    out[MLP_B1 + l] = gb1;
    out[MLP_B2 + l] = gb2;
    if (l < 16) out[MLP_B3 + l] = gb3;
}

// the per-wave partial gradients summed in a fixed order: block b takes
// parameters 64 b .. 64 b + 63 (one per lane), its four waves a quarter of
// the partial sets each (four interleaved sums per lane), and the waves'
// sums are added in wave order
__global__ void __launch_bounds__(256) k_mlp_reduce(int nwaves, const float *__restrict__ partial,
                                                    float *__restrict__ grads) {
    __shared__ float sh[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (p < MLP_NP) {
        const int q = (nwaves + 3) / 4;
        const int w0 = wv * q, w1 = min(nwaves, w0 + q);
        int w = w0;
        for (; w + 3 < w1; w += 4) {
            s0 += partial[(size_t)w * MLP_NP + p];
            s1 += partial[(size_t)(w + 1) * MLP_NP + p];
            s2 += partial[(size_t)(w + 2) * MLP_NP + p];
            s3 += partial[(size_t)(w + 3) * MLP_NP + p];
        }
        for (; w < w1; ++w) s0 += partial[(size_t)w * MLP_NP + p];
    }
    sh[wv][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wv != 0 || p >= MLP_NP) return;
    float t = sh[0][lane] + sh[1][lane] + sh[2][lane] + sh[3][lane];
    // padding entries (W1 columns 8..15, W3 rows 13..15, b3 13..15) stay 0
    if (p >= MLP_W1 && p < MLP_B1 && (p - MLP_W1) % 16 >= MLP_IN) t = 0;
    if (p >= MLP_W3 && p < MLP_B3 && (p - MLP_W3) / 64 >= MLP_OUT) t = 0;
    if (p >= MLP_B3 && p - MLP_B3 >= MLP_OUT) t = 0;
    grads[p] = t;
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp; v[p] = vp;
    w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ---------------------------------------------------------------------
// the probe loss and dL/dM
// ---------------------------------------------------------------------
// Rademacher probe entry of (seed, probe p, cell i)
__host__ __device__ inline float probe_sign(unsigned seed, unsigned p, unsigned i) {
    unsigned h = seed * 0x9E3779B1u ^ p * 0x85EBCA77u ^ i * 0xC2B2AE3Du;
    h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
    return (h >> 16) & 1u ? 1.0f : -1.0f;
}

__global__ void k_probe(int N, unsigned seed, unsigned p, float *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) v[i] = probe_sign(seed, p, (unsigned)i);
}

// u = M v; e = v - A u (A, M: 13 planes over op)
__global__ void k_stencil(StencilOp op, const float *__restrict__ M, const float *__restrict__ v,
                          float *__restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    float s = 0;
    for (int k = 0; k < op.ns; ++k) s += M[(size_t)k * op.N + i] * v[stencil_col(op, ix, iy, k)];
    u[i] = s;
}
__global__ void k_err(StencilOp op, const float *__restrict__ A, const float *__restrict__ u,
                      const float *__restrict__ v, float *__restrict__ e, double *__restrict__ part) {
    double acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < op.N; i += gridDim.x * blockDim.x) {
        const int ix = i % op.nx, iy = i / op.nx;
        float s = 0;
        for (int k = 0; k < op.ns; ++k) s += A[(size_t)k * op.N + i] * u[stencil_col(op, ix, iy, k)];
        const float ei = v[i] - s;
        e[i] = ei;
        acc += (double)ei * ei;
    }
    __shared__ double sh[MBLK / 64];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < MBLK / 64; ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}
// the cells i along one axis whose slot offset o folds onto cell j: the
// direct one j - o, and at a Symmetry edge the mirrored ones -1 - j - o and
// 2n - 1 - j - o (three distinct candidates, each kept when it lies in the
// grid and really folds onto j); at a Periodic edge the one wrap
__device__ inline int fold_sources(int j, int o, int n, int periodic, int *src) {
    if (periodic) { int i = (j - o) % n; if (i < 0) i += n; src[0] = i; return 1; }
    int m = 0;
    const int cand[3] = {j - o, -1 - j - o, 2 * n - 1 - j - o};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int i = cand[c];
        if (i >= 0 && i < n && fold_axis(i + o, n, 0) == j) src[m++] = i;
    }
    return m;
}
// w = A^T e by gathering, in a fixed order, the entries of every row that
// reads column j (deterministic, no atomics)
__global__ void k_transpose_apply(StencilOp op, const float *__restrict__ A, const float *__restrict__ e,
                                  float *__restrict__ w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= op.N) return;
    const int jx = j % op.nx, jy = j / op.nx;
    float s = 0;
    for (int k = 0; k < op.ns; ++k) {
        int sx[3], sy[3];
        const int mx = fold_sources(jx, op.ox[k], op.nx, op.perx, sx);
        const int my = op.dim == 2 ? fold_sources(jy, op.oy[k], op.ny, op.pery, sy) : (sy[0] = 0, 1);
        for (int a = 0; a < my; ++a)
            for (int b = 0; b < mx; ++b) {
                const int i = sy[a] * op.nx + sx[b];
                s += A[(size_t)k * op.N + i] * e[i];
            }
    }
    w[j] = s;
}
// G[k][i] (+)= coef w_i v_col(i,k)
__global__ void k_gacc(StencilOp op, const float *__restrict__ w, const float *__restrict__ v, float coef,
                       int first, float *__restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    const float wi = coef * w[i];
    for (int k = 0; k < op.ns; ++k) {
        const float g = wi * v[stencil_col(op, ix, iy, k)];
        G[(size_t)k * op.N + i] = first ? g : G[(size_t)k * op.N + i] + g;
    }
}
__global__ void k_sum_part(int n, const double *__restrict__ part, double *__restrict__ out, int add) {
    __shared__ double sh[MBLK];
    double s = 0;
    for (int b = threadIdx.x; b < n; b += MBLK) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = MBLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = add ? *out + sh[0] : sh[0];
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static constexpr int LOSS_BLOCKS = 512;

// deterministic initial weights: He-scaled normals from a fixed-seed
// generator, the output layer near zero with the centre bias at 1 (M starts
// as the row scale times the identity)
static void init_params(std::vector<float> &p) {
    p.assign(MLP_NP, 0.0f);
    unsigned long long z = 0x5EED5EEDULL;
    auto uni = [&]() {
        z += 0x9E3779B97F4A7C15ULL;
        unsigned long long x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        x ^= x >> 31;
        return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int o = 0; o < 64; ++o)
        for (int i = 0; i < MLP_IN; ++i) p[MLP_W1 + o * 16 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
    for (int o = 0; o < 64; ++o)
        for (int i = 0; i < 64; ++i) p[MLP_W2 + o * 64 + i] = (float)(normal() * std::sqrt(2.0 / 64));
    for (int o = 0; o < MLP_OUT; ++o)
        for (int i = 0; i < 64; ++i) p[MLP_W3 + o * 64 + i] = (float)(normal() * 0.01 / 8.0);
    p[MLP_B3 + 6] = 1.0f;   // centre slot of the 2-D stencil
}

int mlp_init(Mlp &m, int N, hipStream_t st) {
    if (m.params && m.N == N) return 0;
    mlp_free(m);
    m.N = N;
    m.nwaves = TRAIN_BLOCKS * WAVES;
    auto fa = [](float **p, size_t n) { return hipMalloc((void **)p, n * sizeof(float)) == hipSuccess ? 0 : -1; };
    if (fa(&m.params, MLP_NP) || fa(&m.grads, MLP_NP) || fa(&m.adam_m, MLP_NP) || fa(&m.adam_v, MLP_NP)) return -1;
    if (fa(&m.partial, (size_t)m.nwaves * MLP_NP)) return -1;
    if (fa(&m.G, (size_t)13 * N) || fa(&m.M, (size_t)13 * N)) return -1;
    if (fa(&m.v, N) || fa(&m.u, N) || fa(&m.e, N) || fa(&m.w, N)) return -1;
    MLP_CHECK(hipMalloc((void **)&m.loss_part, LOSS_BLOCKS * sizeof(double)));
    MLP_CHECK(hipMalloc((void **)&m.loss_dev, 2 * sizeof(double)));
    MLP_CHECK(hipHostMalloc((void **)&m.loss_host, 2 * sizeof(double), hipHostMallocDefault));
    std::vector<float> p;
    init_params(p);
    MLP_CHECK(hipMemcpy(m.params, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    MLP_CHECK(hipMemset(m.adam_m, 0, MLP_NP * sizeof(float)));
    MLP_CHECK(hipMemset(m.adam_v, 0, MLP_NP * sizeof(float)));
    m.adam_t = 0;
    (void)st;
    return 0;
}

void mlp_free(Mlp &m) {
    for (float *p : {m.params, m.grads, m.adam_m, m.adam_v, m.partial, m.G, m.M, m.v, m.u, m.e, m.w})
        if (p) (void)hipFree(p);
    if (m.loss_part) (void)hipFree(m.loss_part);
    if (m.loss_dev) (void)hipFree(m.loss_dev);
    if (m.loss_host) (void)hipHostFree(m.loss_host);
    m = Mlp();
}

// blocks of the inference: each block stages the weights once, then its
// waves keep them in registers over many 16-cell tiles
static constexpr int INFER_BLOCKS = 1024;

int mlp_infer(Mlp &m, int N, const float *feat, const float *mscale, float *M, hipStream_t st) {
    const int ntiles = (N + 15) / 16;
    const int blocks = std::min((ntiles + WAVES - 1) / WAVES, INFER_BLOCKS);
    hipLaunchKernelGGL(k_mlp_forward<false>, dim3(blocks), dim3(MBLK), 0, st, N, (const float *)m.params, feat, mscale,
                       M, (__half *)nullptr);
    MLP_CHECK(hipGetLastError());
    return 0;
}

int mlp_infer_half(Mlp &m, int N, const float *feat, __half *Y, hipStream_t st) {
    const int ntiles = (N + 15) / 16;
    const int blocks = std::min((ntiles + WAVES - 1) / WAVES, INFER_BLOCKS);
    hipLaunchKernelGGL(k_mlp_forward<true>, dim3(blocks), dim3(MBLK), 0, st, N, (const float *)m.params, feat,
                       (const float *)nullptr, (float *)nullptr, Y);
    MLP_CHECK(hipGetLastError());
    return 0;
}

int mlp_loss_grad(Mlp &m, const SolverT<float> &S, const float *feat, const float *mscale, int nprobe,
                  unsigned seed, double *loss, hipStream_t st) {
    const StencilOp &op = S.op0;
    const int N = op.N;
    const int nbk = (N + MBLK - 1) / MBLK;
    if (mlp_infer(m, N, feat, mscale, m.M, st)) return -1;
    const float coef = -2.0f / ((float)nprobe * (float)N);
    for (int p = 0; p < nprobe; ++p) {
        hipLaunchKernelGGL(k_probe, dim3(nbk), dim3(MBLK), 0, st, N, seed, (unsigned)p, m.v);
        hipLaunchKernelGGL(k_stencil, dim3(nbk), dim3(MBLK), 0, st, op, (const float *)m.M, (const float *)m.v, m.u);
        hipLaunchKernelGGL(k_err, dim3(LOSS_BLOCKS), dim3(MBLK), 0, st, op, S.A, (const float *)m.u,
                           (const float *)m.v, m.e, m.loss_part);
        hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(MBLK), 0, st, LOSS_BLOCKS, (const double *)m.loss_part,
                           m.loss_dev, p > 0 ? 1 : 0);
        hipLaunchKernelGGL(k_transpose_apply, dim3(nbk), dim3(MBLK), 0, st, op, S.A, (const float *)m.e, m.w);
        hipLaunchKernelGGL(k_gacc, dim3(nbk), dim3(MBLK), 0, st, op, (const float *)m.w, (const float *)m.v, coef,
                           p == 0 ? 1 : 0, m.G);
    }
    hipLaunchKernelGGL(k_mlp_backward, dim3(TRAIN_BLOCKS), dim3(MBLK), 0, st, N, (const float *)m.params, feat,
                       mscale, (const float *)m.G, m.partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + 63) / 64), dim3(256), 0, st, m.nwaves,
                       (const float *)m.partial, m.grads);
    MLP_CHECK(hipGetLastError());
    if (loss) {
        MLP_CHECK(hipMemcpyAsync(m.loss_host, m.loss_dev, sizeof(double), hipMemcpyDeviceToHost, st));
        MLP_CHECK(hipStreamSynchronize(st));
        *loss = m.loss_host[0] / ((double)nprobe * N);
    }
    return 0;
}

int mlp_train(Mlp &m, const SolverT<float> &S, const float *feat, const float *mscale, int iters, double lr,
              int nprobe, unsigned seed, double *loss_first, double *loss_last, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; ++it) {
        double l = 0;
        const bool want = (it == 0) || (it == iters - 1);
        if (mlp_loss_grad(m, S, feat, mscale, nprobe, seed + 7919u * (unsigned)(m.adam_t + 1), want ? &l : nullptr, st))
            return -1;
        if (it == 0 && loss_first) *loss_first = l;
        if (it == iters - 1 && loss_last) *loss_last = l;
        ++m.adam_t;
        const float c1 = 1.0f - std::pow(b1, (float)m.adam_t), c2 = 1.0f - std::pow(b2, (float)m.adam_t);
        hipLaunchKernelGGL(k_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, st, m.params, (const float *)m.grads,
                           m.adam_m, m.adam_v, (float)lr, b1, b2, c1, c2, eps);
        MLP_CHECK(hipGetLastError());
    }
    m.train_iters += iters;
    return 0;
}

}  // namespace yz
