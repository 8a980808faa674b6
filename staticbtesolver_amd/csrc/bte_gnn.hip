// bte_gnn.hip -- mesh graph network for the learned outer-iteration correction: forward, backward
// and Adam on gfx950.  Dense products run on gfx950's BF16 MFMA tiles, v_mfma_f32_16x16x32_bf16,
// with FP32 accumulation (the inference layers fused into one launch each).  Operands are FP32 in
// HBM and rounded to BF16 (nearest even) when loaded into the MFMA registers.  Network and parameter layout: oracle/gnn_ref.py restates them on
// the CPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "../../include/bte.h"
#include "bte_internal.hpp"

namespace bte {

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// gfx950's v_mfma_f32_16x16x32_bf16 operand layout: lane l holds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16] for j = 0..7; the result element acc[i] is C[4 (l / 16) + i][l % 16].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short short8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 pack_bf16x8(const float (&v)[8]) {
    short8 s;
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] = f2bf(v[j]);
    return __builtin_bit_cast(bf16x8, s);
}

// One wave accumulates a 16 x 64 output block: acc[t] is the 16 x 16 tile of columns 16t..16t+15.
// a_at(r, k): row r (0..15) of the A block, column k; b_at(k, n): B[k][n].  K is a multiple of 16;
// the k steps are 32 wide, a last half step padded with zeros.
template <class FA, class FB>
__device__ __forceinline__ void wave_mm_16x64(const FA& a_at, const FB& b_at, int K, f32x4 (&acc)[4]) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bool in = k0 + kq < K;  // a lane's 8 k values are all in range or all past K
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; j++) av[j] = in ? a_at(r, k0 + kq + j) : 0.f;
        const bf16x8 a = pack_bf16x8(av);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float bv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) bv[j] = in ? b_at(k0 + kq + j, 16 * t + r) : 0.f;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pack_bf16x8(bv), acc[t], 0, 0, 0);
        }
    }
}

#define GNN_ROW0 ((blockIdx.x * 4 + (threadIdx.x >> 6)) * 16)

// Encoder: Z0 = X W_in + b_in, H0 = relu(Z0).  X is [N][16], W_in [16][64].
__global__ __launch_bounds__(256) void k_gnn_enc(int N, const float* __restrict__ X, const float* __restrict__ W,
                                                const float* __restrict__ b, float* __restrict__ Z,
                                                float* __restrict__ H) {
    const int row0 = GNN_ROW0;
    if (row0 >= N) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    f32x4 acc[4] = {};
    wave_mm_16x64([&](int r, int k) { return (row0 + r < N) ? X[(size_t)(row0 + r) * kGnnIn + k] : 0.f; },
                  [&](int k, int n) { return W[k * kGnnHid + n]; }, kGnnIn, acc);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = row0 + 4 * (lane >> 4) + i, col = 16 * t + (lane & 15);
            if (row < N) {
                const float z = acc[t][i] + b[col];
                if (Z) Z[(size_t)row * kGnnHid + col] = z;  // null: inference keeps no pre-activation
                H[(size_t)row * kGnnHid + col] = fmaxf(z, 0.f);
            }
        }
}

// Message-passing layer: Z = H W1 + M W2 + b, Hn = H + relu(Z).
__global__ __launch_bounds__(256) void k_gnn_layer(int N, const float* __restrict__ H, const float* __restrict__ M,
                                                  const float* __restrict__ W1, const float* __restrict__ W2,
                                                  const float* __restrict__ b, float* __restrict__ Z,
                                                  float* __restrict__ Hn) {
    const int row0 = GNN_ROW0;
    if (row0 >= N) return;
    const int lane = threadIdx.x & 63;
    f32x4 acc[4] = {};
    wave_mm_16x64([&](int r, int k) { return (row0 + r < N) ? H[(size_t)(row0 + r) * kGnnHid + k] : 0.f; },
                  [&](int k, int n) { return W1[k * kGnnHid + n]; }, kGnnHid, acc);
    wave_mm_16x64([&](int r, int k) { return (row0 + r < N) ? M[(size_t)(row0 + r) * kGnnHid + k] : 0.f; },
                  [&](int k, int n) { return W2[k * kGnnHid + n]; }, kGnnHid, acc);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = row0 + 4 * (lane >> 4) + i, col = 16 * t + (lane & 15);
            if (row < N) {
                const size_t o = (size_t)row * kGnnHid + col;
                const float z = acc[t][i] + b[col];
                Z[o] = z;
                Hn[o] = H[o] + fmaxf(z, 0.f);
            }
        }
}

// Input gradient of a layer product: G = dZ W^T (W is [64][64]).
__global__ __launch_bounds__(256) void k_gnn_dgrad(int N, const float* __restrict__ dZ, const float* __restrict__ W,
                                                  float* __restrict__ G) {
    const int row0 = GNN_ROW0;
    if (row0 >= N) return;
    const int lane = threadIdx.x & 63;
    f32x4 acc[4] = {};
    wave_mm_16x64([&](int r, int k) { return (row0 + r < N) ? dZ[(size_t)(row0 + r) * kGnnHid + k] : 0.f; },
                  [&](int k, int n) { return W[n * kGnnHid + k]; }, kGnnHid, acc);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = row0 + 4 * (lane >> 4) + i, col = 16 * t + (lane & 15);
            if (row < N) G[(size_t)row * kGnnHid + col] = acc[t][i];
        }
}

// Weight gradient A^T G over a chunk of kGnnChunk nodes per work-group: partial[(i * 64 + j) * nblk
// + blk].  A is [N][KA] (KA = 16 or 64), G is [N][64]; wave w computes rows i in [16w, 16w + 16).
constexpr int kGnnChunk = 512;
__global__ __launch_bounds__(256) void k_gnn_wgrad(int N, int KA, const float* __restrict__ A,
                                                  const float* __restrict__ G, float* __restrict__ partial) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = 16 * wave;
    if (i0 >= KA) return;  // uniform per wave
    const int n_beg = blockIdx.x * kGnnChunk;
    const int n_end = min(N, n_beg + kGnnChunk);
    f32x4 acc[4] = {};
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int n0 = n_beg; n0 < n_end; n0 += 32) {
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n = n0 + kq + j;
            av[j] = n < n_end ? A[(size_t)n * KA + i0 + r] : 0.f;
        }
        const bf16x8 a = pack_bf16x8(av);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float bv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int n = n0 + kq + j;
                bv[j] = n < n_end ? G[(size_t)n * kGnnHid + 16 * t + r] : 0.f;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pack_bf16x8(bv), acc[t], 0, 0, 0);
        }
    }
    const int nblk = gridDim.x;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int i = i0 + 4 * (lane >> 4) + e, j = 16 * t + (lane & 15);
            partial[(size_t)(i * kGnnHid + j) * nblk + blockIdx.x] = acc[t][e];
        }
}

// Column sums sum_c A[c][k] * s[c] (s == nullptr: 1) per block: partial[k * nblk + blk].
__global__ __launch_bounds__(256) void k_gnn_colsum(int N, int C, const float* __restrict__ A,
                                                   const float* __restrict__ s, float* __restrict__ partial) {
    // 256 threads = 4 node lanes x 64 columns; a block covers kGnnChunk nodes
    const int k = threadIdx.x & 63, q = threadIdx.x >> 6;
    __shared__ float red[4][64];
    float acc = 0.f;
    if (k < C) {
        const int n_beg = blockIdx.x * kGnnChunk, n_end = min(N, n_beg + kGnnChunk);
        for (int n = n_beg + q; n < n_end; n += 4) acc += A[(size_t)n * C + k] * (s ? s[n] : 1.f);
    }
    red[q][k] = acc;
    __syncthreads();
    if (q == 0 && k < C) partial[(size_t)k * gridDim.x + blockIdx.x] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// out[q] = sum_b partial[q * nblk + b] (fixed order)
__global__ void k_gnn_reduce(int nq, int nblk, const float* __restrict__ partial, float* __restrict__ out) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int b = 0; b < nblk; b++) s += partial[(size_t)q * nblk + b];
        out[q] = s;
    }
}

// M[c] = mean of H over the interior neighbours of c.
__global__ void k_gnn_agg(int N, int F, const int* __restrict__ nbr, const float* __restrict__ H,
                          float* __restrict__ M) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)N * kGnnHid;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t / kGnnHid), k = (int)(t % kGnnHid);
        float s = 0.f;
        int d = 0;
        for (int f = 0; f < F; f++) {
            const int j = nbr[(size_t)c * F + f];
            if (j >= 0) {
                s += H[(size_t)j * kGnnHid + k];
                d++;
            }
        }
        M[t] = d ? s / (float)d : 0.f;
    }
}

// dH_out[c] = dH[c] + G1[c] + sum_{j in nb(c)} G2[j] / deg(j)   (transpose of the mean aggregation)
__global__ void k_gnn_dh(int N, int F, const int* __restrict__ nbr, const int* __restrict__ deg,
                         const float* __restrict__ dH, const float* __restrict__ G1, const float* __restrict__ G2,
                         float* __restrict__ out) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)N * kGnnHid;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t / kGnnHid), k = (int)(t % kGnnHid);
        float s = 0.f;
        for (int f = 0; f < F; f++) {
            const int j = nbr[(size_t)c * F + f];
            if (j >= 0) s += G2[(size_t)j * kGnnHid + k] / (float)deg[j];
        }
        out[t] = dH[t] + G1[t] + s;
    }
}

__global__ void k_gnn_deg(int N, int F, const int* __restrict__ nbr, int* __restrict__ deg) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        int d = 0;
        for (int f = 0; f < F; f++) d += nbr[(size_t)c * F + f] >= 0;
        deg[c] = d;
    }
}

// Decoder y = H w_out + b_out (one thread per node, columns summed in order).
__global__ void k_gnn_dec(int N, const float* __restrict__ H, const float* __restrict__ w, const float* __restrict__ b,
                          float* __restrict__ y) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int k = 0; k < kGnnHid; k++) s += H[(size_t)c * kGnnHid + k] * w[k];
        y[c] = s + b[0];
    }
}

// dy = 2 w (y - t) / sw;  per-node loss terms in lt
__global__ void k_gnn_loss(int N, const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ w,
                           float inv_sw, float* __restrict__ dy, float* __restrict__ lt) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        const float wc = w ? w[c] : 1.f, e = y[c] - t[c];
        dy[c] = 2.f * wc * e * inv_sw;
        lt[c] = wc * e * e * inv_sw;
    }
}

// dH = dy (x) w_out
__global__ void k_gnn_dec_bwd(int N, const float* __restrict__ dy, const float* __restrict__ w, float* __restrict__ dH) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)N * kGnnHid;
         t += (long long)gridDim.x * blockDim.x)
        dH[t] = dy[t / kGnnHid] * w[t % kGnnHid];
}

// dZ = dH where Z > 0
__global__ void k_gnn_relu_bwd(long long n, const float* __restrict__ Z, const float* __restrict__ dH,
                               float* __restrict__ dZ) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dZ[t] = Z[t] > 0.f ? dH[t] : 0.f;
}

__global__ void k_gnn_adam(int P, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// ---- inference: fused layers --------------------------------------------------------------------

// Operand fragments of the message-passing weights, [layer][matrix (W1, W2)][k block][n tile][lane]:
// lane l of fragment (kb, t) holds W[32 kb + 8 (l / 16) + j][16 t + l % 16], j = 0..7, as BF16.
__global__ void k_gnn_wfrag(int L, const float* __restrict__ p, bf16x8* __restrict__ frag) {
    const int total = L * 2 * 2 * 4 * 64;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int lane = e & 63, t = (e >> 6) & 3, kb = (e >> 8) & 1, m = (e >> 9) & 1, l = e >> 10;
        const float* W = p + gnn_off_layer(l) + m * kGnnHid * kGnnHid;
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = W[(32 * kb + 8 * (lane >> 4) + j) * kGnnHid + 16 * t + (lane & 15)];
        frag[e] = pack_bf16x8(v);
    }
}

// One message-passing layer for inference, H_out = H + relu(H W1 + M W2 + b) with M the mean of H
// over the interior neighbours (as k_gnn_agg forms it, in FP32 and the same order), without storing
// M or the pre-activation.  Each wave takes 16-row tiles; its 16 weight fragments stay in
// registers.  The tile's own rows pass through the wave's LDS tile twice: in the operand layout
// for the residual term read in the result layout, then H_out in the result layout for stores of
// whole rows.  Waves take different numbers of tiles, so only wave-scope ordering is used.  Work-group
// blk runs on XCD blk % 8, and each XCD takes a contiguous range of tiles (the grid is a multiple of
// 8), so a tile's neighbour rows, mostly in nearby tiles, come from that XCD's L2.  The F <= FM
// neighbour indices are loaded together, then every neighbour row (an absent one adds 0 in its
// place in the sum): two memory round trips instead of two per neighbour.  DEC: the last layer --
// instead of H_out, the decoder y = H_out w_out + b_out.
template <bool DEC, int FM>
__global__ __launch_bounds__(256) void k_gnn_infer_layer(int N, int F, const int* __restrict__ nbr,
                                                        const float* __restrict__ H, const bf16x8* __restrict__ wf,
                                                        const float* __restrict__ b, float* __restrict__ Hn,
                                                        const float* __restrict__ wo, const float* __restrict__ bo,
                                                        float* __restrict__ y) {
    constexpr int TS = kGnnHid + 4;  // LDS row stride (floats): rows stay 16-byte aligned
    __shared__ float tiles[4][16 * TS];
    float* tl = tiles[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int ntile = (N + 15) / 16;
    bf16x8 w1[2][4], w2[2][4];
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            w1[kb][t] = wf[((0 * 2 + kb) * 4 + t) * 64 + lane];
            w2[kb][t] = wf[((1 * 2 + kb) * 4 + t) * 64 + lane];
        }
    float bias[4], wout[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        bias[t] = b[16 * t + r];
        wout[t] = DEC ? wo[16 * t + r] : 0.f;
    }
    const int per = gridDim.x / 8;  // work-groups per XCD
    const int lblk = (blockIdx.x % 8) * per + blockIdx.x / 8;
    for (int tile = lblk * 4 + (threadIdx.x >> 6); tile < ntile; tile += gridDim.x * 4) {
        // A operands: row tile*16 + r (clamped: rows past N compute discarded outputs), columns
        // 32 kb + 8 q + j of H and of the neighbour mean
        const int rr = min(tile * 16 + r, N - 1);
        float h[2][8], mv[2][8];
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            const float4* src = (const float4*)(H + (size_t)rr * kGnnHid + 32 * kb + 8 * q);
            const float4 u0 = src[0], u1 = src[1];
            h[kb][0] = u0.x, h[kb][1] = u0.y, h[kb][2] = u0.z, h[kb][3] = u0.w;
            h[kb][4] = u1.x, h[kb][5] = u1.y, h[kb][6] = u1.z, h[kb][7] = u1.w;
            float4* dst = (float4*)(tl + r * TS + 32 * kb + 8 * q);
            dst[0] = u0;
            dst[1] = u1;
#pragma unroll
            for (int j = 0; j < 8; j++) mv[kb][j] = 0.f;
        }
        int nbf[FM];
#pragma unroll
        for (int f = 0; f < FM; f++) nbf[f] = f < F ? nbr[(size_t)rr * F + f] : -1;
        float4 u[FM][2][2];
#pragma unroll
        for (int f = 0; f < FM; f++)
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const float4* src = (const float4*)(H + (size_t)max(nbf[f], 0) * kGnnHid + 32 * kb + 8 * q);
                u[f][kb][0] = src[0];
                u[f][kb][1] = src[1];
            }
        int d = 0;
#pragma unroll
        for (int f = 0; f < FM; f++) {
            if (nbf[f] < 0) continue;
            d++;
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const float4 u0 = u[f][kb][0], u1 = u[f][kb][1];
                mv[kb][0] += u0.x, mv[kb][1] += u0.y, mv[kb][2] += u0.z, mv[kb][3] += u0.w;
                mv[kb][4] += u1.x, mv[kb][5] += u1.y, mv[kb][6] += u1.z, mv[kb][7] += u1.w;
            }
        }
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int j = 0; j < 8; j++) mv[kb][j] = d ? mv[kb][j] / (float)d : 0.f;
        f32x4 acc[4] = {};
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            const bf16x8 a = pack_bf16x8(h[kb]);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w1[kb][t], acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            const bf16x8 a = pack_bf16x8(mv[kb]);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w2[kb][t], acc[t], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the tile's rows are in LDS
        // result element (row 4 q + i, column 16 t + r)
        float hn[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int t = 0; t < 4; t++) hn[i][t] = tl[(4 * q + i) * TS + 16 * t + r] + fmaxf(acc[t][i] + bias[t], 0.f);
        if (DEC) {
            // y of row 4 q + i: the 64 column terms over the 16 lanes of the group (r = 0..15)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float v = 0.f;
#pragma unroll
                for (int t = 0; t < 4; t++) v += hn[i][t] * wout[t];
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
                const int row = tile * 16 + 4 * q + i;
                if (r == 0 && row < N) y[row] = v + bo[0];
            }
        } else {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // every lane has read its residuals
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int t = 0; t < 4; t++) tl[(4 * q + i) * TS + 16 * t + r] = hn[i][t];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // the 16 rows are contiguous in H_out: 256 float4, four per lane
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = lane + 64 * k, row = e >> 4, c4 = e & 15;
                if (tile * 16 + row < N)
                    *(float4*)(Hn + (size_t)(tile * 16 + row) * kGnnHid + 4 * c4) = *(const float4*)(tl + row * TS + 4 * c4);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // LDS reads done before the next tile's writes
    }
}

static int grid1(long long n, int b = 256) { return (int)std::max<long long>(1, std::min<long long>((n + b - 1) / b, 65536)); }

int gnn_ensure(GnnNet* net, int n) {
    if (n <= net->cap && net->H.size() == (size_t)net->L + 1 && net->Z0.p) return 0;
    net->cap = 0;
    const size_t a = sizeof(float) * (size_t)n * kGnnHid;
    net->H.clear();
    net->M.clear();
    net->Z.clear();
    net->H.resize(net->L + 1);
    net->M.resize(net->L);
    net->Z.resize(net->L);
    for (auto& b : net->H)
        if (b.ensure(a) != hipSuccess) return BTE_ERR_HIP;
    for (auto& b : net->M)
        if (b.ensure(a) != hipSuccess) return BTE_ERR_HIP;
    for (auto& b : net->Z)
        if (b.ensure(a) != hipSuccess) return BTE_ERR_HIP;
    for (DBuf* b : {&net->Z0, &net->dH, &net->dH2, &net->dZ, &net->G1, &net->G2})
        if (b->ensure(a) != hipSuccess) return BTE_ERR_HIP;
    for (DBuf* b : {&net->y, &net->tgt, &net->wgt, &net->dy})
        if (b->ensure(sizeof(float) * n) != hipSuccess) return BTE_ERR_HIP;
    if (net->X.ensure(sizeof(float) * (size_t)n * kGnnIn) != hipSuccess) return BTE_ERR_HIP;
    const int nblk = grid1(n, kGnnChunk);
    net->part_elems = (size_t)kGnnHid * kGnnHid * nblk + (size_t)n;
    if (net->part.ensure(sizeof(float) * net->part_elems) != hipSuccess) return BTE_ERR_HIP;
    if (net->red.ensure(sizeof(float) * (kGnnHid * kGnnHid + 64)) != hipSuccess) return BTE_ERR_HIP;
    net->cap = n;
    return 0;
}

int gnn_forward_device(GnnNet* net, hipStream_t stream) {
    const int N = net->N, L = net->L;
    const float* p = net->params.as<float>();
    const int gt = (N + 63) / 64;  // 64 rows per work-group
    hipStream_t st = stream ? stream : net->stream;
    hipLaunchKernelGGL(k_gnn_enc, dim3(gt), dim3(256), 0, st, N, net->X.as<float>(), p, p + kGnnIn * kGnnHid,
                       net->Z0.as<float>(), net->H[0].as<float>());
    for (int l = 0; l < L; l++) {
        const float* W1 = p + gnn_off_layer(l);
        hipLaunchKernelGGL(k_gnn_agg, dim3(grid1((long long)N * kGnnHid)), dim3(256), 0, st, N, net->F,
                           net->nbr.as<int>(), net->H[l].as<float>(), net->M[l].as<float>());
        hipLaunchKernelGGL(k_gnn_layer, dim3(gt), dim3(256), 0, st, N, net->H[l].as<float>(), net->M[l].as<float>(), W1,
                           W1 + kGnnHid * kGnnHid, W1 + 2 * kGnnHid * kGnnHid, net->Z[l].as<float>(),
                           net->H[l + 1].as<float>());
    }
    const float* wo = p + gnn_off_out(L);
    hipLaunchKernelGGL(k_gnn_dec, dim3(grid1(N)), dim3(256), 0, st, N, net->H[L].as<float>(), wo, wo + kGnnHid,
                       net->y.as<float>());
    return hipGetLastError() == hipSuccess ? 0 : BTE_ERR_HIP;
}

int gnn_infer_device(GnnNet* net, hipStream_t stream) {
    const int N = net->N, L = net->L;
    const float* p = net->params.as<float>();
    hipStream_t st = stream ? stream : net->stream;
    const size_t per_layer = 2 * 2 * 4 * 64;
    if (net->wfrag_ver != net->params_ver) {
        if (net->wfrag.ensure(sizeof(bf16x8) * per_layer * L) != hipSuccess) return BTE_ERR_HIP;
        hipLaunchKernelGGL(k_gnn_wfrag, dim3(grid1((long long)per_layer * L)), dim3(256), 0, st, L, p,
                           net->wfrag.as<bf16x8>());
        net->wfrag_ver = net->params_ver;
    }
    hipLaunchKernelGGL(k_gnn_enc, dim3((N + 63) / 64), dim3(256), 0, st, N, net->X.as<float>(), p,
                       p + kGnnIn * kGnnHid, (float*)nullptr, net->H[0].as<float>());
    const int ntile = (N + 15) / 16;
    const dim3 grid(std::min(((ntile + 3) / 4 + 7) / 8 * 8, 65536));  // a tile per wave, a multiple of 8
    const float* wo = p + gnn_off_out(L);
    for (int l = 0; l < L; l++) {
        const float* H = net->H[l & 1].as<float>();
        float* Hn = net->H[(l + 1) & 1].as<float>();
        const bf16x8* wf = net->wfrag.as<bf16x8>() + per_layer * l;
        const float* b = p + gnn_off_layer(l) + 2 * kGnnHid * kGnnHid;
        // the faces per cell of 1D, 2D and 3D meshes (2, 3, 4) unrolled exactly, others (up to 8) as 8
        auto go = [&](auto dec) {
            constexpr bool D = decltype(dec)::value;
            auto launch = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, grid, dim3(256), 0, st, N, net->F, net->nbr.as<int>(), H, wf, b, Hn, wo,
                                   wo + kGnnHid, net->y.as<float>());
            };
            switch (net->F) {
                case 2: launch(k_gnn_infer_layer<D, 2>); break;
                case 3: launch(k_gnn_infer_layer<D, 3>); break;
                case 4: launch(k_gnn_infer_layer<D, 4>); break;
                default: launch(k_gnn_infer_layer<D, 8>); break;
            }
        };
        if (l + 1 < L)
            go(std::false_type{});
        else
            go(std::true_type{});
    }
    return hipGetLastError() == hipSuccess ? 0 : BTE_ERR_HIP;
}

// Weight gradient A^T G -> out[KA][64] (reduced over node chunks).
static int wgrad(GnnNet* net, int KA, const float* A, const float* G, float* out) {
    const int nblk = grid1(net->N, kGnnChunk);
    hipLaunchKernelGGL(k_gnn_wgrad, dim3(nblk), dim3(256), 0, net->stream, net->N, KA, A, G, net->part.as<float>());
    hipLaunchKernelGGL(k_gnn_reduce, dim3(grid1(KA * kGnnHid)), dim3(256), 0, net->stream, KA * kGnnHid, nblk,
                       net->part.as<float>(), out);
    return hipGetLastError() == hipSuccess ? 0 : BTE_ERR_HIP;
}

static int colsum(GnnNet* net, int C, const float* A, const float* s, float* out) {
    const int nblk = grid1(net->N, kGnnChunk);
    hipLaunchKernelGGL(k_gnn_colsum, dim3(nblk), dim3(256), 0, net->stream, net->N, C, A, s, net->part.as<float>());
    hipLaunchKernelGGL(k_gnn_reduce, dim3(1), dim3(64), 0, net->stream, C, nblk, net->part.as<float>(), out);
    return hipGetLastError() == hipSuccess ? 0 : BTE_ERR_HIP;
}

// Backward pass after gnn_forward_device: loss and gradients into net->grads.
static int gnn_backward_device(GnnNet* net, float inv_sw, bool weighted, double* loss) {
    const int N = net->N, L = net->L;
    const float* p = net->params.as<float>();
    float* g = net->grads.as<float>();
    const long long NH = (long long)N * kGnnHid;
    const int gt = (N + 63) / 64;
    float* lt = net->part.as<float>() + (net->part_elems - N);  // per-node loss terms (tail of part)
    hipLaunchKernelGGL(k_gnn_loss, dim3(grid1(N)), dim3(256), 0, net->stream, N, net->y.as<float>(), net->tgt.as<float>(),
                       weighted ? net->wgt.as<float>() : nullptr, inv_sw, net->dy.as<float>(), lt);
    std::vector<float> hl(N);
    if (hipMemcpyAsync(hl.data(), lt, sizeof(float) * N, hipMemcpyDeviceToHost, net->stream) != hipSuccess)
        return BTE_ERR_HIP;
    const float* wo = p + gnn_off_out(L);
    int rc;
    if ((rc = colsum(net, kGnnHid, net->H[L].as<float>(), net->dy.as<float>(), g + gnn_off_out(L)))) return rc;
    if ((rc = colsum(net, 1, net->dy.as<float>(), nullptr, g + gnn_off_out(L) + kGnnHid))) return rc;
    hipLaunchKernelGGL(k_gnn_dec_bwd, dim3(grid1(NH)), dim3(256), 0, net->stream, N, net->dy.as<float>(), wo,
                       net->dH.as<float>());
    float* dH = net->dH.as<float>();
    float* dH2 = net->dH2.as<float>();
    for (int l = L - 1; l >= 0; l--) {
        const int off = gnn_off_layer(l);
        hipLaunchKernelGGL(k_gnn_relu_bwd, dim3(grid1(NH)), dim3(256), 0, net->stream, NH, net->Z[l].as<float>(), dH,
                           net->dZ.as<float>());
        if ((rc = wgrad(net, kGnnHid, net->H[l].as<float>(), net->dZ.as<float>(), g + off))) return rc;
        if ((rc = wgrad(net, kGnnHid, net->M[l].as<float>(), net->dZ.as<float>(), g + off + kGnnHid * kGnnHid)))
            return rc;
        if ((rc = colsum(net, kGnnHid, net->dZ.as<float>(), nullptr, g + off + 2 * kGnnHid * kGnnHid))) return rc;
        hipLaunchKernelGGL(k_gnn_dgrad, dim3(gt), dim3(256), 0, net->stream, N, net->dZ.as<float>(), p + off,
                           net->G1.as<float>());
        hipLaunchKernelGGL(k_gnn_dgrad, dim3(gt), dim3(256), 0, net->stream, N, net->dZ.as<float>(),
                           p + off + kGnnHid * kGnnHid, net->G2.as<float>());
        hipLaunchKernelGGL(k_gnn_dh, dim3(grid1(NH)), dim3(256), 0, net->stream, N, net->F, net->nbr.as<int>(),
                           net->deg.as<int>(), dH, net->G1.as<float>(), net->G2.as<float>(), dH2);
        std::swap(dH, dH2);
    }
    hipLaunchKernelGGL(k_gnn_relu_bwd, dim3(grid1(NH)), dim3(256), 0, net->stream, NH, net->Z0.as<float>(), dH,
                       net->dZ.as<float>());
    if ((rc = wgrad(net, kGnnIn, net->X.as<float>(), net->dZ.as<float>(), g))) return rc;
    if ((rc = colsum(net, kGnnHid, net->dZ.as<float>(), nullptr, g + kGnnIn * kGnnHid))) return rc;
    if (hipStreamSynchronize(net->stream) != hipSuccess) return BTE_ERR_HIP;
    double s = 0;
    for (float v : hl) s += v;
    if (loss) *loss = s;
    return 0;
}

}  // namespace bte

using namespace bte;

#define GCHK(expr)                                                                   \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            g->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
            return BTE_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

extern "C" {

int bte_gnn_create(int device, int layers, unsigned long long seed, bte_gnn** out) {
    if (!out || layers < 1 || layers > 8) return BTE_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return BTE_ERR_NODEVICE;
    bte_gnn* g = new bte_gnn();
    if (device < 0) (void)hipGetDevice(&device);
    g->device = device;
    g->L = layers;
    g->P = gnn_num_params(layers);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        g->params.ensure(sizeof(float) * g->P) != hipSuccess || g->grads.ensure(sizeof(float) * g->P) != hipSuccess ||
        g->adam_m.ensure(sizeof(float) * g->P) != hipSuccess || g->adam_v.ensure(sizeof(float) * g->P) != hipSuccess) {
        delete g;
        return BTE_ERR_HIP;
    }
    // He-style initialisation of the weight blocks, zero biases, small output layer
    std::vector<float> p(g->P, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int count, double sd) {
        std::normal_distribution<double> nd(0.0, sd);
        for (int i = 0; i < count; i++) p[off + i] = (float)nd(rng);
    };
    fill(0, kGnnIn * kGnnHid, std::sqrt(2.0 / kGnnIn));
    for (int l = 0; l < layers; l++) {
        fill(gnn_off_layer(l), 2 * kGnnHid * kGnnHid, 0.5 / std::sqrt((double)kGnnHid));
    }
    fill(gnn_off_out(layers), kGnnHid, 0.01);
    (void)hipMemcpy(g->params.p, p.data(), sizeof(float) * g->P, hipMemcpyHostToDevice);
    g->params_ver++;
    (void)hipMemset(g->adam_m.p, 0, sizeof(float) * g->P);
    (void)hipMemset(g->adam_v.p, 0, sizeof(float) * g->P);
    (void)hipMemset(g->grads.p, 0, sizeof(float) * g->P);
    *out = g;
    return BTE_OK;
}

void bte_gnn_destroy(bte_gnn* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) {
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamDestroy(g->stream);
    }
    delete g;
}

const char* bte_gnn_last_error(const bte_gnn* g) { return g ? g->err.c_str() : "null network"; }

int bte_gnn_num_params(const bte_gnn* g) { return g ? g->P : 0; }

int bte_gnn_get_params(bte_gnn* g, float* params) {
    if (!g || !params) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    GCHK(hipMemcpy(params, g->params.p, sizeof(float) * g->P, hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_gnn_set_params(bte_gnn* g, const float* params) {
    if (!g || !params) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    GCHK(hipMemcpy(g->params.p, params, sizeof(float) * g->P, hipMemcpyHostToDevice));
    g->params_ver++;
    GCHK(hipMemset(g->adam_m.p, 0, sizeof(float) * g->P));
    GCHK(hipMemset(g->adam_v.p, 0, sizeof(float) * g->P));
    g->adam_t = 0;
    return BTE_OK;
}

int bte_gnn_get_grads(bte_gnn* g, float* grads) {
    if (!g || !grads) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    GCHK(hipMemcpy(grads, g->grads.p, sizeof(float) * g->P, hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_gnn_set_graph(bte_gnn* g, int n_node, int n_face, const int* nbr) {
    if (!g || n_node <= 0 || n_face <= 0 || n_face > 8 || !nbr) return BTE_ERR_ARG;
    for (long long i = 0; i < (long long)n_node * n_face; i++)
        if (nbr[i] >= n_node) {
            g->err = "bte_gnn_set_graph: neighbour index out of range";
            return BTE_ERR_ARG;
        }
    GCHK(hipSetDevice(g->device));
    g->N = 0;
    g->graph_tag = 0;
    if (gnn_ensure(g, n_node)) {
        g->err = "out of device memory";
        return BTE_ERR_HIP;
    }
    g->N = n_node;
    g->F = n_face;
    GCHK(g->nbr.ensure(sizeof(int) * (size_t)n_node * n_face));
    GCHK(g->deg.ensure(sizeof(int) * (size_t)n_node));
    GCHK(hipMemcpyAsync(g->nbr.p, nbr, sizeof(int) * (size_t)n_node * n_face, hipMemcpyHostToDevice, g->stream));
    hipLaunchKernelGGL(k_gnn_deg, dim3(grid1(n_node)), dim3(256), 0, g->stream, n_node, n_face, g->nbr.as<int>(),
                       g->deg.as<int>());
    GCHK(hipGetLastError());
    GCHK(hipStreamSynchronize(g->stream));
    return BTE_OK;
}

int bte_gnn_forward(bte_gnn* g, const float* x, float* y) {
    if (!g || !x || g->N <= 0) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    GCHK(hipMemcpyAsync(g->X.p, x, sizeof(float) * (size_t)g->N * kGnnIn, hipMemcpyHostToDevice, g->stream));
    if (gnn_infer_device(g)) {
        g->err = "forward launch failed";
        return BTE_ERR_HIP;
    }
    if (y) GCHK(hipMemcpyAsync(y, g->y.p, sizeof(float) * g->N, hipMemcpyDeviceToHost, g->stream));
    GCHK(hipStreamSynchronize(g->stream));
    return BTE_OK;
}

int bte_gnn_backward(bte_gnn* g, const float* x, const float* target, const float* weight, double* loss) {
    if (!g || !x || !target || g->N <= 0) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    const int N = g->N;
    double sw = N;
    if (weight) {
        sw = 0;
        for (int i = 0; i < N; i++) sw += weight[i];
        if (!(sw > 0)) {
            g->err = "node weights must have a positive sum";
            return BTE_ERR_ARG;
        }
        GCHK(hipMemcpyAsync(g->wgt.p, weight, sizeof(float) * N, hipMemcpyHostToDevice, g->stream));
    }
    GCHK(hipMemcpyAsync(g->X.p, x, sizeof(float) * (size_t)N * kGnnIn, hipMemcpyHostToDevice, g->stream));
    GCHK(hipMemcpyAsync(g->tgt.p, target, sizeof(float) * N, hipMemcpyHostToDevice, g->stream));
    if (gnn_forward_device(g) || gnn_backward_device(g, (float)(1.0 / sw), weight != nullptr, loss)) {
        g->err = "backward launch failed";
        return BTE_ERR_HIP;
    }
    return BTE_OK;
}

int bte_gnn_adam(bte_gnn* g, float lr, float beta1, float beta2, float eps, float weight_decay) {
    if (!g) return BTE_ERR_ARG;
    GCHK(hipSetDevice(g->device));
    g->adam_t++;
    g->params_ver++;
    const float c1 = 1.f - std::pow(beta1, (float)g->adam_t), c2 = 1.f - std::pow(beta2, (float)g->adam_t);
    hipLaunchKernelGGL(k_gnn_adam, dim3(grid1(g->P)), dim3(256), 0, g->stream, g->P, g->params.as<float>(),
                       g->grads.as<float>(), g->adam_m.as<float>(), g->adam_v.as<float>(), lr, beta1, beta2, eps,
                       weight_decay, c1, c2);
    GCHK(hipGetLastError());
    GCHK(hipStreamSynchronize(g->stream));
    return BTE_OK;
}

}  // extern "C"
