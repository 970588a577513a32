// mlp.hip - the learned cut proposal: a point-set network that ranks a part's
// candidate planes, trained on exact scores the library computes itself.
//
// No reference counterpart. Network (PointNet-style, conditioned on the
// plane): 256 area-stratified surface samples of a part in the part's box
// frame (center, half the largest extent); for a candidate plane (unit normal
// n, offset d in the same frame) every sample gets its signed distance
// s = n.p + d, and (x, y, z, s) goes through a shared point MLP
// 4 -> 128 -> 128 -> 128 (ReLU); the features are max-pooled separately over
// the samples on each side of the plane (g+, g-: the two halves' shapes) and a
// head [g+, g-, n, d] -> 128 -> 128 -> 1 (ReLU between) predicts the plane's
// exact cut score. BF16 operands, FP32 accumulation: every layer is one
// call of a wave64 MFMA GEMM (v_mfma_f32_16x16x16_bf16) with a fused epilogue
// (bias, ReLU, backward ReLU mask, FP32 and BF16 stores). Training keeps FP32
// master weights and Adam state on the device and runs forward, backward and
// the update as kernels (mean squared error on the score).
//
// Parameter layout (FP32, padded to multiples of 16; the padding stays zero):
//   W1 [16 x 128] (rows 0..3 used), b1 [128], W2 [128 x 128], b2, W3, b3,
//   H1 [272 x 128] (rows 0..259 used), c1 [128], H2 [128 x 128], c2,
//   H3 [128 x 16] (column 0 used), c3 [16] (entry 0 used).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "kerr.h"
#include "mlp.hpp"
#include "wave.hpp"

// signed distances and sums round as written (the oracle repeats them)
#pragma clang fp contract(off)

using namespace cuacd;

namespace {

typedef unsigned short bf16_t;
typedef short s4v_t __attribute__((ext_vector_type(4)));
typedef float f4v_t __attribute__((ext_vector_type(4)));

constexpr int NS = MLP_SAMPLES;  // samples per part
constexpr int HD = 128;   // hidden width
constexpr int KIN = 16;   // padded point input width
constexpr int KH = 272;   // padded head input width (2 x 128 + 4)
constexpr int NO = 16;    // padded output width

// offsets of the parameter blocks (floats)
constexpr int P_W1 = 0, P_B1 = P_W1 + KIN * HD, P_W2 = P_B1 + HD, P_B2 = P_W2 + HD * HD, P_W3 = P_B2 + HD,
              P_B3 = P_W3 + HD * HD, P_H1 = P_B3 + HD, P_C1 = P_H1 + KH * HD, P_H2 = P_C1 + HD,
              P_C2 = P_H2 + HD * HD, P_H3 = P_C2 + HD, P_C3 = P_H3 + HD * NO, P_TOTAL = P_C3 + NO;

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(bf16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// GEMM: C[M x N] = alpha * A[M x K] . B[K x N] (+ bias[n]) (ReLU) (* (mask > 0))
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn] (BF16), so
// transposed operands are strides. 64 x 64 output tile per 256-thread
// workgroup, K in steps of 32 staged through LDS; each wave owns a 32 x 32
// quarter as 2 x 2 MFMA 16 x 16 x 16 tiles. Operand layout of the MFMA (lane
// l): A row l % 16, k 4 (l / 16) .. +3; B column l % 16, same k; D column
// l % 16, rows 4 (l / 16) .. +3.
// ---------------------------------------------------------------------------
struct GemmArgs {
    int M, N, K;
    const bf16_t* A;
    long long sam, sak;
    const bf16_t* B;
    long long sbk, sbn;
    float alpha;
    const float* bias;   // [N] or null
    int relu;
    const bf16_t* mask;  // [M x N] with row stride ldm, or null: out *= (mask > 0)
    long long ldm;
    float* C;            // FP32 out (row stride ldc) or null
    long long ldc;
    bf16_t* Cb;          // BF16 out (row stride ldcb) or null
    long long ldcb;
    int kchunk;          // split K: workgroup z covers k in [z kchunk, (z + 1) kchunk)
    float* part;         // split K: raw partial sums [gridDim.z x M x N] (epilogue in the reduction)
};

constexpr int GM = 64, GN = 64, GK = 32, GKP = GK + 8;  // LDS rows padded by 16 bytes

__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g) {
    __shared__ __align__(16) bf16_t As[GM][GKP];
    __shared__ __align__(16) bf16_t Bt[GN][GKP];  // transposed: Bt[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.y * GM, n0 = blockIdx.x * GN;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    f4v_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4v_t{0.f, 0.f, 0.f, 0.f};
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // loader: one row (or column), 8 consecutive k
    const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
    for (int k0 = kb; k0 < ke; k0 += GK) {
        {
            const int m = m0 + lr;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + lk + j;
                As[lr][lk + j] = (m < g.M && k < ke) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : (bf16_t)0;
            }
            const int n = n0 + lr;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + lk + j;
                Bt[lr][lk + j] = (n < g.N && k < ke) ? g.B[(long long)k * g.sbk + (long long)n * g.sbn] : (bf16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 16) {
            const int kk = ks + 4 * (lane >> 4);
            s4v_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const bf16_t* pa = &As[wm + 16 * i + (lane & 15)][kk];
                a[i] = s4v_t{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                const bf16_t* pb = &Bt[wn + 16 * i + (lane & 15)][kk];
                b[i] = s4v_t{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + wn + 16 * j + (lane & 15);
            if (n >= g.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                if (m >= g.M) continue;
                if (g.part) {
                    g.part[((long long)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][r];
                    continue;
                }
                float v = g.alpha * acc[i][j][r];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = fmaxf(v, 0.0f);
                if (g.mask && !(bf2f(g.mask[(long long)m * g.ldm + n]) > 0.0f)) v = 0.0f;
                if (g.C) g.C[(long long)m * g.ldc + n] = v;
                if (g.Cb) g.Cb[(long long)m * g.ldcb + n] = f2bf(v);
            }
        }
}

// ---------------------------------------------------------------------------
// Fused inference of the point layers and the pooling: a persistent grid of
// 256-thread workgroups, each with the transposed point-layer weights in LDS,
// strides over the candidates; per candidate wave w owns samples 64w..64w+63
// (4 x 8 MFMA 16 x 16 x 16 tiles) and the activations stay in LDS as BF16.
// The MFMA operand layout, the k order (16-wide steps, ascending; layer 1
// adds the GEMM's second, all-zero step), the epilogue (bias, ReLU, BF16
// rounding) and the pooling rule are those of mlp_gemm_kernel and
// mlp_maxpool, so G equals the layer-by-layer path's bit for bit; only
// G [C x 256] reaches memory.
// ---------------------------------------------------------------------------
constexpr int FP = HD + 8;    // activation / weight row pitch in LDS (bf16)
constexpr int FP1 = KIN + 8;  // pitch of the transposed first layer
constexpr unsigned FUSED_LDS = (unsigned)(NS * FP * 2 + 2 * HD * FP * 2 + HD * FP1 * 2 + NS * 4 + 2 * 4 * HD * 4);

__device__ __forceinline__ s4v_t lds_frag(const bf16_t* p) { return *reinterpret_cast<const s4v_t*>(p); }

__global__ __launch_bounds__(256) void mlp_encode_pool_kernel(const float* pts, const float* planes,
                                                              const int* part_of, int C, const bf16_t* W,
                                                              const float* Wf, float* G) {
    extern __shared__ __align__(16) char fl[];
    bf16_t* H = (bf16_t*)fl;                    // [NS][FP] X, then the layer outputs
    bf16_t* W2t = H + NS * FP;                  // [HD][FP] W2 transposed: W2t[n][k]
    bf16_t* W3t = W2t + HD * FP;                // [HD][FP]
    bf16_t* W1t = W3t + HD * FP;                // [HD][FP1]
    float* Sd = (float*)(W1t + HD * FP1);       // [NS] signed distances
    float* red = Sd + NS;                       // [2][4][HD] per-wave side maxima
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < HD * HD; i += 256) {
        const int k = i / HD, n = i % HD;
        W2t[n * FP + k] = W[P_W2 + i];
        W3t[n * FP + k] = W[P_W3 + i];
    }
    for (int i = tid; i < KIN * HD; i += 256) {
        const int k = i / HD, n = i % HD;
        W1t[n * FP1 + k] = W[P_W1 + i];
    }
    const int r0 = 64 * wv;
    const int fr = lane & 15, fk = 4 * (lane >> 4);
    for (int c = blockIdx.x; c < C; c += gridDim.x) {
        __syncthreads();
        {
            // sample tid of the candidate's part (the rows of mlp_pack_points)
            const float* pp = pts + ((long long)part_of[c] * NS + tid) * 3;
            const float* pl = planes + 4ll * c;
            const float sd = pl[0] * pp[0] + pl[1] * pp[1] + pl[2] * pp[2] + pl[3];
            Sd[tid] = sd;
            bf16_t* x = H + tid * FP;
            x[0] = f2bf(pp[0]);
            x[1] = f2bf(pp[1]);
            x[2] = f2bf(pp[2]);
            x[3] = f2bf(sd);
            for (int j = 4; j < KIN; ++j) x[j] = 0;
        }
        __syncthreads();
        f4v_t acc[4][8];
        // layer 1: K = 16
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = f4v_t{0.f, 0.f, 0.f, 0.f};
        {
            s4v_t a[4], b[8];
            const s4v_t z = s4v_t{0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = lds_frag(H + (r0 + 16 * i + fr) * FP + fk);
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = lds_frag(W1t + (16 * j + fr) * FP1 + fk);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(z, z, acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
        auto store = [&](const float* bias) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int n = 16 * j + fr;
                    const float bn = bias[n];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float v = 1.0f * acc[i][j][r];
                        v += bn;
                        v = fmaxf(v, 0.0f);
                        H[(r0 + 16 * i + fk + r) * FP + n] = f2bf(v);
                    }
                }
        };
        auto layer = [&](const bf16_t* Wt) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = f4v_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int k0 = 0; k0 < HD; k0 += 16) {
                s4v_t a[4], b[8];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = lds_frag(H + (r0 + 16 * i + fr) * FP + k0 + fk);
#pragma unroll
                for (int j = 0; j < 8; ++j) b[j] = lds_frag(Wt + (16 * j + fr) * FP + k0 + fk);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        };
        store(Wf + P_B1);
        __syncthreads();
        layer(W2t);
        __syncthreads();
        store(Wf + P_B2);
        __syncthreads();
        layer(W3t);
        // layer 3 epilogue: BF16-rounded activations pooled per side of the plane
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int n = 16 * j + fr;
            const float bn = Wf[P_B3 + n];
            float bp = -1e30f, bq = -1e30f;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float v = 1.0f * acc[i][j][r];
                    v += bn;
                    v = fmaxf(v, 0.0f);
                    v = bf2f(f2bf(v));
                    if (Sd[r0 + 16 * i + fk + r] >= 0.0f) bp = fmaxf(bp, v);
                    else bq = fmaxf(bq, v);
                }
            bp = fmaxf(bp, __shfl_xor(bp, 16));
            bp = fmaxf(bp, __shfl_xor(bp, 32));
            bq = fmaxf(bq, __shfl_xor(bq, 16));
            bq = fmaxf(bq, __shfl_xor(bq, 32));
            if (lane < 16) {
                red[(0 * 4 + wv) * HD + n] = bp;
                red[(1 * 4 + wv) * HD + n] = bq;
            }
        }
        __syncthreads();
        {
            const int side = tid >> 7, n = tid & (HD - 1);
            float m = red[(side * 4 + 0) * HD + n];
#pragma unroll
            for (int q = 1; q < 4; ++q) m = fmaxf(m, red[(side * 4 + q) * HD + n]);
            G[(long long)c * 2 * HD + tid] = m > -1e30f ? m : 0.0f;
        }
    }
}

// Split-K reduction with the GEMM's epilogue: C = alpha * sum_z part[z] ...
__global__ void mlp_splitk_reduce(GemmArgs g, int nz) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.M * g.N) return;
    const int m = (int)(i / g.N), n = (int)(i % g.N);
    float v = 0.0f;
    for (int z = 0; z < nz; ++z) v += g.part[(long long)z * g.M * g.N + i];
    v *= g.alpha;
    if (g.bias) v += g.bias[n];
    if (g.relu) v = fmaxf(v, 0.0f);
    if (g.mask && !(bf2f(g.mask[(long long)m * g.ldm + n]) > 0.0f)) v = 0.0f;
    if (g.C) g.C[(long long)m * g.ldc + n] = v;
    if (g.Cb) g.Cb[(long long)m * g.ldcb + n] = f2bf(v);
}

// out[n] = sum over rows of D [R x N] (BF16): per block a 1024-row chunk of
// 128 columns into part[chunk][n], then a sum over the chunks
__global__ __launch_bounds__(128) void mlp_colsum_part(const bf16_t* D, int R, int N, float* part) {
    const int n = blockIdx.x * 128 + threadIdx.x, r0 = blockIdx.y * 1024;
    if (n >= N) return;
    float s = 0.0f;
    const int r1 = min(R, r0 + 1024);
    for (int r = r0; r < r1; ++r) s += bf2f(D[(long long)r * N + n]);
    part[(long long)blockIdx.y * N + n] = s;
}
__global__ void mlp_colsum_fin(const float* part, int chunks, int N, float* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[(long long)c * N + n];
    out[n] = s;
}

// ---------------------------------------------------------------------------
// Elementwise / reduction kernels
// ---------------------------------------------------------------------------
// X [C*NS x 16] BF16: candidate c's rows are its part's samples with their
// signed distance to the plane (x, y, z, s, zeros); S [C*NS] the distances
__global__ void mlp_pack_points(const float* pts, const float* planes, const int* part_of, int C, bf16_t* X,
                                float* S) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long long)C * NS) return;
    const long long c = r / NS;
    const int i = (int)(r % NS);
    const float* p = pts + ((long long)part_of[c] * NS + i) * 3;
    const float* pl = planes + 4 * c;
    const float sd = pl[0] * p[0] + pl[1] * p[1] + pl[2] * p[2] + pl[3];
    S[r] = sd;
    bf16_t* x = X + r * KIN;
    x[0] = f2bf(p[0]);
    x[1] = f2bf(p[1]);
    x[2] = f2bf(p[2]);
    x[3] = f2bf(sd);
    for (int j = 4; j < KIN; ++j) x[j] = 0;
}

// per candidate and channel: max of A [C*NS x 128] over its rows with s >= 0
// (G[c][j]) and with s < 0 (G[c][128 + j]); arg = the row, -1 (and 0) when a
// side has no sample
__global__ __launch_bounds__(HD) void mlp_maxpool(const bf16_t* A, const float* S, int C, float* G, int* arg) {
    const int c = blockIdx.x, j = threadIdx.x;
    if (c >= C) return;
    float bp = -1e30f, bn = -1e30f;
    int ip = -1, in = -1;
    for (int i = 0; i < NS; ++i) {
        const long long r = (long long)c * NS + i;
        const float v = bf2f(A[r * HD + j]);
        if (S[r] >= 0.0f) {
            if (v > bp) { bp = v; ip = (int)r; }
        } else if (v > bn) {
            bn = v;
            in = (int)r;
        }
    }
    G[(long long)c * 2 * HD + j] = ip >= 0 ? bp : 0.0f;
    G[(long long)c * 2 * HD + HD + j] = in >= 0 ? bn : 0.0f;
    arg[(long long)c * 2 * HD + j] = ip;
    arg[(long long)c * 2 * HD + HD + j] = in;
}

// head input [C x 272] BF16: g+, g-, plane (n, d), zeros
__global__ void mlp_head_input(const float* G, const float* planes, int C, bf16_t* Hin) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)C * KH) return;
    const long long c = i / KH;
    const int j = (int)(i % KH);
    float v = 0.0f;
    if (j < 2 * HD) v = G[c * 2 * HD + j];
    else if (j < 2 * HD + 4) v = planes[4 * c + (j - 2 * HD)];
    Hin[i] = f2bf(v);
}

// loss = mean (y - t)^2 over the C candidates; dY [C x 16] BF16 (column 0)
__global__ __launch_bounds__(256) void mlp_loss(const float* Y, const float* t, int C, bf16_t* dY, float* loss) {
    __shared__ float red[4];
    float s = 0.0f;
    for (int c = threadIdx.x; c < C; c += 256) {
        const float e = Y[(long long)c * NO] - t[c];
        s += e * e;
        for (int j = 0; j < NO; ++j) dY[(long long)c * NO + j] = j == 0 ? f2bf(2.0f * e / (float)C) : (bf16_t)0;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) loss[0] = (red[0] + red[1] + red[2] + red[3]) / (float)C;
}


// dZ3 [C*NS x 128] BF16: each pooled gradient dHin[c][0, 256) back to its
// max row, where the max was positive (ReLU)
__global__ void mlp_unpool(const float* dHin, const int* arg, const bf16_t* A3, int C, bf16_t* dZ) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long R = (long long)C * NS;
    if (i >= R * HD) return;
    const long long r = i / HD;
    const int j = (int)(i % HD);
    const long long c = r / NS;
    float v = 0.0f;
    if (bf2f(A3[i]) > 0.0f) {
        if (arg[c * 2 * HD + j] == r) v += dHin[c * KH + j];
        if (arg[c * 2 * HD + HD + j] == r) v += dHin[c * KH + HD + j];
    }
    dZ[i] = f2bf(v);
}

// Adam (FP32 master weights), then the BF16 copy used by the GEMMs
__global__ void mlp_adam(float* w, const float* gr, float* m, float* v, bf16_t* wb, int n, float lr, float b1,
                         float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] -= upd;
    wb[i] = f2bf(w[i]);
}

__global__ void mlp_to_bf16(const float* w, bf16_t* wb, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// ---------------------------------------------------------------------------
// Surface samples of a part: NS area-stratified points in the part's box
// frame. One workgroup per part; triangle areas are prefix-summed in chunks
// into `cum` (nt floats per part at cum_off[p]). Sample i takes the triangle
// holding u = (i + 1/2) / NS of the total area and a point of it from a hash
// of (part, i).
// ---------------------------------------------------------------------------
struct SampleTask {
    const float* v;
    const int* t;
    int nv, nt;
    long long cum_off;
    float lo[3], hi[3];  // part box (for the frame)
};

__device__ __forceinline__ uint32_t mlp_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void mlp_sample_kernel(const SampleTask* tasks, int P, float* cum_all, float* out) {
    __shared__ float wsum[4];
    __shared__ float carry;
    const int p = blockIdx.x, tid = threadIdx.x;
    if (p >= P) return;
    const SampleTask T = tasks[p];
    float* cum = cum_all + T.cum_off;
    if (tid == 0) carry = 0.0f;
    __syncthreads();
    for (int b = 0; b < T.nt; b += 256) {
        const int i = b + tid;
        float a = 0.0f;
        if (i < T.nt) {
            const int i0 = T.t[3 * i], i1 = T.t[3 * i + 1], i2 = T.t[3 * i + 2];
            const float ux = T.v[3 * i1] - T.v[3 * i0], uy = T.v[3 * i1 + 1] - T.v[3 * i0 + 1],
                        uz = T.v[3 * i1 + 2] - T.v[3 * i0 + 2];
            const float wx = T.v[3 * i2] - T.v[3 * i0], wy = T.v[3 * i2 + 1] - T.v[3 * i0 + 1],
                        wz = T.v[3 * i2 + 2] - T.v[3 * i0 + 2];
            const float cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
            a = 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);
        }
        float s = a;  // inclusive wave scan
        for (int o = 1; o < 64; o <<= 1) {
            const float y = __shfl_up(s, o);
            if ((tid & 63) >= o) s += y;
        }
        if ((tid & 63) == 63) wsum[tid >> 6] = s;
        __syncthreads();
        float base = carry;
        for (int q = 0; q < (tid >> 6); ++q) base += wsum[q];
        if (i < T.nt) cum[i] = base + s;
        __syncthreads();
        if (tid == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    const float total = carry;
    float c[3], sc = 0.0f;
    for (int j = 0; j < 3; ++j) {
        c[j] = 0.5f * (T.lo[j] + T.hi[j]);
        sc = fmaxf(sc, 0.5f * (T.hi[j] - T.lo[j]));
    }
    const float inv = sc > 0.0f ? 1.0f / sc : 0.0f;
    for (int i = tid; i < NS; i += 256) {
        float px = 0.f, py = 0.f, pz = 0.f;
        if (T.nt > 0 && total > 0.0f) {
            const float u = ((float)i + 0.5f) / (float)NS * total;
            int lo = 0, hi = T.nt - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cum[mid] > u) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t h1 = mlp_hash((uint32_t)p * 0x9e3779b1u + (uint32_t)i * 0x85ebca6bu + 1u);
            const uint32_t h2 = mlp_hash(h1 ^ 0x68e31da4u);
            float r1 = (float)(h1 >> 8) * (1.0f / 16777216.0f), r2 = (float)(h2 >> 8) * (1.0f / 16777216.0f);
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            const int i0 = T.t[3 * lo], i1 = T.t[3 * lo + 1], i2 = T.t[3 * lo + 2];
            const float* a = T.v + 3 * i0;
            const float* b = T.v + 3 * i1;
            const float* d = T.v + 3 * i2;
            px = a[0] + r1 * (b[0] - a[0]) + r2 * (d[0] - a[0]);
            py = a[1] + r1 * (b[1] - a[1]) + r2 * (d[1] - a[1]);
            pz = a[2] + r1 * (b[2] - a[2]) + r2 * (d[2] - a[2]);
        }
        float* o = out + ((long long)p * NS + i) * 3;
        o[0] = (px - c[0]) * inv;
        o[1] = (py - c[1]) * inv;
        o[2] = (pz - c[2]) * inv;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct gpu_mlp {
    float* w;      // FP32 master weights [P_TOTAL]
    bf16_t* wb;    // BF16 copy
    float* g;      // gradients
    float* m;      // Adam moments
    float* v;
    int step;
    // grow-only activations and split-K / column-sum partials
    char* act;
    size_t act_cap;
    float* part;
    size_t part_cap;
};

namespace {

template <class T> T* carve(char*& p, size_t n) {
    T* r = (T*)p;
    p += (n * sizeof(T) + 255) / 256 * 256;
    return r;
}

// split-K workspace of at least n floats (grow-only; a growth waits for the
// stream so no launch still reads the old buffer)
float* part_ws(gpu_ctx_t ctx, gpu_mlp* mp, size_t n) {
    if (n > mp->part_cap) {
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
        if (mp->part) (void)hipFree(mp->part);
        mp->part = nullptr;
        mp->part_cap = 0;
        if (hipMalloc((void**)&mp->part, n * sizeof(float)) != hipSuccess) return nullptr;
        mp->part_cap = n;
    }
    return mp->part;
}

int gemm(gpu_ctx_t ctx, gpu_mlp* mp, int M, int N, int K, const bf16_t* A, long long sam, long long sak, const bf16_t* B,
         long long sbk, long long sbn, const float* bias, int relu, const bf16_t* mask, long long ldm, float* C,
         long long ldc, bf16_t* Cb, long long ldcb, float alpha = 1.0f) {
    if (M <= 0 || N <= 0) return 0;
    GemmArgs g{M, N, K, A, sam, sak, B, sbk, sbn, alpha, bias, relu, mask, ldm, C, ldc, Cb, ldcb, K, nullptr};
    const int tm = (M + GM - 1) / GM, tn = (N + GN - 1) / GN;
    // few output tiles over a long K (the weight gradients): split K across
    // workgroups, partial sums reduced by a second kernel with the epilogue
    int nz = 1;
    if (tm * tn < 128 && K >= 4096) {
        nz = std::min((1024 + tm * tn - 1) / (tm * tn), (K + 511) / 512);
        const int kc = ((K + nz - 1) / nz + GK - 1) / GK * GK;
        nz = (K + kc - 1) / kc;
        g.kchunk = kc;
        g.part = part_ws(ctx, mp, (size_t)nz * M * N);
        REQUIRE(g.part, "mlp: split-K workspace allocation failed");
    }
    hipLaunchKernelGGL(mlp_gemm_kernel, dim3((unsigned)tn, (unsigned)tm, (unsigned)nz), dim3(256), 0, ctx->stream, g);
    HIP_TRY(hipGetLastError());
    if (nz > 1) {
        const long long n = (long long)M * N;
        hipLaunchKernelGGL(mlp_splitk_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, g, nz);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

struct Act {
    bf16_t *X, *A1, *A2, *A3, *Hin, *B1, *B2, *dY, *dZ2h, *dZ1h, *dA3, *dZ2e, *dZ1e;
    float *S, *G, *Y, *dHin, *loss, *csum;
    size_t csum_cap;
    int* arg;
};

// activations of C candidates (C * NS encoder rows); `train` adds the
// backward buffers
int act_layout(gpu_ctx_t ctx, gpu_mlp* mp, int C, bool train, Act* a, bool fused = false) {
    // the fused inference keeps the encoder rows (X, A1-A3, S) in LDS
    const size_t R = fused ? 0 : (size_t)C * NS;
    size_t need = 0;
    auto sz = [&](size_t n, size_t e) { need += (n * e + 255) / 256 * 256; };
    sz(R * KIN, 2); sz(R * HD, 2); sz(R * HD, 2); sz(R * HD, 2);
    sz((size_t)C * KH, 2); sz((size_t)C * HD, 2); sz((size_t)C * HD, 2); sz((size_t)C * NO, 2);
    if (train) { sz((size_t)C * HD, 2); sz((size_t)C * HD, 2); sz(R * HD, 2); sz(R * HD, 2); sz(R * HD, 2); }
    sz(R, 4); sz((size_t)C * 2 * HD, 4); sz((size_t)C * NO, 4); sz((size_t)C * KH, 4); sz(64, 4);
    sz((size_t)C * 2 * HD, 4);
    const size_t csum = ((R + 1023) / 1024 + 1) * (size_t)HD;
    if (train) sz(csum, 4);
    if (need > mp->act_cap) {
        if (mp->act) HIP_TRY(hipFree(mp->act));
        mp->act = nullptr;
        mp->act_cap = 0;
        HIP_TRY(hipMalloc((void**)&mp->act, need));
        mp->act_cap = need;
    }
    char* p = mp->act;
    *a = Act{};
    a->X = carve<bf16_t>(p, R * KIN);
    a->A1 = carve<bf16_t>(p, R * HD);
    a->A2 = carve<bf16_t>(p, R * HD);
    a->A3 = carve<bf16_t>(p, R * HD);
    a->Hin = carve<bf16_t>(p, (size_t)C * KH);
    a->B1 = carve<bf16_t>(p, (size_t)C * HD);
    a->B2 = carve<bf16_t>(p, (size_t)C * HD);
    a->dY = carve<bf16_t>(p, (size_t)C * NO);
    if (train) {
        a->dZ2h = carve<bf16_t>(p, (size_t)C * HD);
        a->dZ1h = carve<bf16_t>(p, (size_t)C * HD);
        a->dA3 = carve<bf16_t>(p, R * HD);
        a->dZ2e = carve<bf16_t>(p, R * HD);
        a->dZ1e = carve<bf16_t>(p, R * HD);
    }
    a->S = carve<float>(p, R);
    a->G = carve<float>(p, (size_t)C * 2 * HD);
    a->Y = carve<float>(p, (size_t)C * NO);
    a->dHin = carve<float>(p, (size_t)C * KH);
    a->loss = carve<float>(p, 64);
    a->arg = carve<int>(p, (size_t)C * 2 * HD);
    if (train) {
        a->csum = carve<float>(p, csum);
        a->csum_cap = csum;
    }
    return 0;
}

// forward over device inputs: points [P*NS x 3], planes [C x 4], part_of [C]
int forward(gpu_ctx_t ctx, gpu_mlp* mp, const float* d_pts, const float* d_planes, const int* d_part, int C,
            Act& a) {
    const int R = C * NS;
    const bf16_t* W = mp->wb;
    const float* Wf = mp->w;
    hipLaunchKernelGGL(mlp_pack_points, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, ctx->stream, d_pts,
                       d_planes, d_part, C, a.X, a.S);
    HIP_TRY(hipGetLastError());
    if (int rc = gemm(ctx, mp, R, HD, KIN, a.X, KIN, 1, W + P_W1, HD, 1, Wf + P_B1, 1, nullptr, 0, nullptr, 0, a.A1, HD))
        return rc;
    if (int rc = gemm(ctx, mp, R, HD, HD, a.A1, HD, 1, W + P_W2, HD, 1, Wf + P_B2, 1, nullptr, 0, nullptr, 0, a.A2, HD))
        return rc;
    if (int rc = gemm(ctx, mp, R, HD, HD, a.A2, HD, 1, W + P_W3, HD, 1, Wf + P_B3, 1, nullptr, 0, nullptr, 0, a.A3, HD))
        return rc;
    hipLaunchKernelGGL(mlp_maxpool, dim3((unsigned)C), dim3(HD), 0, ctx->stream, a.A3, a.S, C, a.G, a.arg);
    HIP_TRY(hipGetLastError());
    {
        const long long n = (long long)C * KH;
        hipLaunchKernelGGL(mlp_head_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a.G,
                           d_planes, C, a.Hin);
        HIP_TRY(hipGetLastError());
    }
    if (int rc = gemm(ctx, mp, C, HD, KH, a.Hin, KH, 1, W + P_H1, HD, 1, Wf + P_C1, 1, nullptr, 0, nullptr, 0, a.B1, HD))
        return rc;
    if (int rc = gemm(ctx, mp, C, HD, HD, a.B1, HD, 1, W + P_H2, HD, 1, Wf + P_C2, 1, nullptr, 0, nullptr, 0, a.B2, HD))
        return rc;
    return gemm(ctx, mp, C, NO, HD, a.B2, HD, 1, W + P_H3, NO, 1, Wf + P_C3, 0, nullptr, 0, a.Y, NO, nullptr, 0);
}

// forward through the fused point-layer kernel (inference only): G, then the head
int forward_fused(gpu_ctx_t ctx, gpu_mlp* mp, const float* d_pts, const float* d_planes, const int* d_part, int C,
                  Act& a) {
    const bf16_t* W = mp->wb;
    const float* Wf = mp->w;
    HIP_TRY(hipFuncSetAttribute((const void*)mlp_encode_pool_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)FUSED_LDS));
    const int grid = std::max(1, std::min(C, ctx->n_cu));
    hipLaunchKernelGGL(mlp_encode_pool_kernel, dim3((unsigned)grid), dim3(256), FUSED_LDS, ctx->stream, d_pts,
                       d_planes, d_part, C, W, Wf, a.G);
    HIP_TRY(hipGetLastError());
    {
        const long long n = (long long)C * KH;
        hipLaunchKernelGGL(mlp_head_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a.G,
                           d_planes, C, a.Hin);
        HIP_TRY(hipGetLastError());
    }
    if (int rc = gemm(ctx, mp, C, HD, KH, a.Hin, KH, 1, W + P_H1, HD, 1, Wf + P_C1, 1, nullptr, 0, nullptr, 0, a.B1, HD))
        return rc;
    if (int rc = gemm(ctx, mp, C, HD, HD, a.B1, HD, 1, W + P_H2, HD, 1, Wf + P_C2, 1, nullptr, 0, nullptr, 0, a.B2, HD))
        return rc;
    return gemm(ctx, mp, C, NO, HD, a.B2, HD, 1, W + P_H3, NO, 1, Wf + P_C3, 0, nullptr, 0, a.Y, NO, nullptr, 0);
}

// gradients of the mean squared error into mp->g (after forward)
int backward(gpu_ctx_t ctx, gpu_mlp* mp, const float* d_t, int C, Act& a) {
    const int R = C * NS;
    // column sums use the part of the activation block reserved for them
    float* colsum_ws = a.csum;
    const size_t colsum_cap = a.csum_cap;
    const bf16_t* W = mp->wb;
    float* g = mp->g;
    hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(256), 0, ctx->stream, a.Y, d_t, C, a.dY, a.loss);
    HIP_TRY(hipGetLastError());
    auto colsum = [&](const bf16_t* D, int rows, int N, float* out) -> int {
        const int chunks = (rows + 1023) / 1024;
        float* part = colsum_ws;
        REQUIRE(part && (size_t)chunks * N <= colsum_cap, "mlp: column-sum workspace too small");
        hipLaunchKernelGGL(mlp_colsum_part, dim3((unsigned)((N + 127) / 128), (unsigned)chunks), dim3(128), 0,
                           ctx->stream, D, rows, N, part);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(mlp_colsum_fin, dim3((unsigned)((N + 127) / 128)), dim3(128), 0, ctx->stream, part, chunks,
                           N, out);
        HIP_TRY(hipGetLastError());
        return 0;
    };
    // head
    if (int rc = gemm(ctx, mp, HD, NO, C, a.B2, 1, HD, a.dY, NO, 1, nullptr, 0, nullptr, 0, g + P_H3, NO, nullptr, 0))
        return rc;
    if (int rc = colsum(a.dY, C, NO, g + P_C3)) return rc;
    if (int rc = gemm(ctx, mp, C, HD, NO, a.dY, NO, 1, W + P_H3, 1, NO, nullptr, 0, a.B2, HD, nullptr, 0, a.dZ2h, HD))
        return rc;
    if (int rc = gemm(ctx, mp, HD, HD, C, a.B1, 1, HD, a.dZ2h, HD, 1, nullptr, 0, nullptr, 0, g + P_H2, HD, nullptr, 0))
        return rc;
    if (int rc = colsum(a.dZ2h, C, HD, g + P_C2)) return rc;
    if (int rc = gemm(ctx, mp, C, HD, HD, a.dZ2h, HD, 1, W + P_H2, 1, HD, nullptr, 0, a.B1, HD, nullptr, 0, a.dZ1h, HD))
        return rc;
    if (int rc = gemm(ctx, mp, KH, HD, C, a.Hin, 1, KH, a.dZ1h, HD, 1, nullptr, 0, nullptr, 0, g + P_H1, HD, nullptr, 0))
        return rc;
    if (int rc = colsum(a.dZ1h, C, HD, g + P_C1)) return rc;
    if (int rc = gemm(ctx, mp, C, KH, HD, a.dZ1h, HD, 1, W + P_H1, 1, HD, nullptr, 0, nullptr, 0, a.dHin, KH, nullptr, 0))
        return rc;
    // encoder
    {
        const long long n = (long long)R * HD;
        hipLaunchKernelGGL(mlp_unpool, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a.dHin, a.arg,
                           a.A3, C, a.dA3);
        HIP_TRY(hipGetLastError());
    }
    if (int rc = gemm(ctx, mp, HD, HD, R, a.A2, 1, HD, a.dA3, HD, 1, nullptr, 0, nullptr, 0, g + P_W3, HD, nullptr, 0))
        return rc;
    if (int rc = colsum(a.dA3, R, HD, g + P_B3)) return rc;
    if (int rc = gemm(ctx, mp, R, HD, HD, a.dA3, HD, 1, W + P_W3, 1, HD, nullptr, 0, a.A2, HD, nullptr, 0, a.dZ2e, HD))
        return rc;
    if (int rc = gemm(ctx, mp, HD, HD, R, a.A1, 1, HD, a.dZ2e, HD, 1, nullptr, 0, nullptr, 0, g + P_W2, HD, nullptr, 0))
        return rc;
    if (int rc = colsum(a.dZ2e, R, HD, g + P_B2)) return rc;
    if (int rc = gemm(ctx, mp, R, HD, HD, a.dZ2e, HD, 1, W + P_W2, 1, HD, nullptr, 0, a.A1, HD, nullptr, 0, a.dZ1e, HD))
        return rc;
    if (int rc = gemm(ctx, mp, KIN, HD, R, a.X, 1, KIN, a.dZ1e, HD, 1, nullptr, 0, nullptr, 0, g + P_W1, HD, nullptr, 0))
        return rc;
    return colsum(a.dZ1e, R, HD, g + P_B1);
}

// Inputs staged on the device (points [P*NS*3], planes [C*4], part_of, cand
// offsets, targets), grow-only per context call.
struct Staged {
    float* pts = nullptr;
    float* planes = nullptr;
    int* part = nullptr;
    int* coff = nullptr;
    float* t = nullptr;
    char* base = nullptr;
    gpu_ctx_t ctx = nullptr;
    ~Staged() {
        if (base) (void)hipFreeAsync(base, ctx->stream);
    }
    int load(gpu_ctx_t c, const float* h_pts, int P, const float* h_planes, const int* h_part, int C,
             const float* h_t, std::vector<int>& coff_h) {
        ctx = c;
        coff_h.assign(P + 1, 0);
        for (int i = 0; i < C; ++i) {
            if (h_part[i] < 0 || h_part[i] >= P) return -2;
            if (i > 0 && h_part[i] < h_part[i - 1]) return -3;
            coff_h[h_part[i] + 1]++;
        }
        for (int p = 0; p < P; ++p) coff_h[p + 1] += coff_h[p];
        const size_t b0 = ((size_t)P * NS * 12 + 255) / 256 * 256, b1 = ((size_t)C * 16 + 255) / 256 * 256,
                     b2 = ((size_t)C * 4 + 255) / 256 * 256, b3 = ((size_t)(P + 1) * 4 + 255) / 256 * 256,
                     b4 = ((size_t)C * 4 + 255) / 256 * 256;
        if (hipMallocAsync((void**)&base, b0 + b1 + b2 + b3 + b4 + 256, c->stream) != hipSuccess) return -1;
        pts = (float*)base;
        planes = (float*)(base + b0);
        part = (int*)(base + b0 + b1);
        coff = (int*)(base + b0 + b1 + b2);
        t = (float*)(base + b0 + b1 + b2 + b3);
        if (hipMemcpyAsync(pts, h_pts, (size_t)P * NS * 12, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(planes, h_planes, (size_t)C * 16, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(part, h_part, (size_t)C * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(coff, coff_h.data(), (size_t)(P + 1) * 4, hipMemcpyHostToDevice, c->stream) !=
                hipSuccess)
            return -1;
        if (h_t && hipMemcpyAsync(t, h_t, (size_t)C * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) return -1;
        return 0;
    }
};

int stage_err(gpu_ctx_t ctx, int rc) {
    if (rc == -2) set_error(ctx, "mlp: part_of index out of range");
    else if (rc == -3) set_error(ctx, "mlp: candidates must be grouped by part (part_of ascending)");
    else set_error(ctx, "mlp: input staging failed");
    return -1;
}

}  // namespace

namespace cuacd {
int mlp_forward_dev(gpu_ctx_t ctx, gpu_mlp* mp, const float* d_pts, int P, const float* d_planes, const int* d_part,
                    int C, float* d_out) {
    // bounded activations: slices of at most 8192 candidates (2M encoder
    // rows) layer by layer; the fused kernel keeps those rows in LDS
    (void)P;
    const bool fused = mlp_fused_on() != 0;
    const int slice = fused ? 65536 : 8192;
    Act a;
    if (int rc = act_layout(ctx, mp, std::min(slice, C), false, &a, fused)) return rc;
    for (int c0 = 0; c0 < C; c0 += slice) {
        const int n = std::min(slice, C - c0);
        if (int rc = fused ? forward_fused(ctx, mp, d_pts, d_planes + 4ll * c0, d_part + c0, n, a)
                           : forward(ctx, mp, d_pts, d_planes + 4ll * c0, d_part + c0, n, a))
            return rc;
        HIP_TRY(hipMemcpy2DAsync(d_out + c0, 4, a.Y, 4 * NO, 4, n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return 0;
}

int mlp_sample_dev(gpu_ctx_t ctx, const MlpSamplePart* parts, int P, float* d_out) {
    if (P <= 0) return 0;
    std::vector<SampleTask> st(P);
    long long off = 0;
    for (int p = 0; p < P; ++p) {
        st[p].v = parts[p].v;
        st[p].t = parts[p].t;
        st[p].nv = parts[p].nv;
        st[p].nt = parts[p].nt;
        st[p].cum_off = off;
        for (int j = 0; j < 3; ++j) { st[p].lo[j] = parts[p].lo[j]; st[p].hi[j] = parts[p].hi[j]; }
        off += parts[p].nt;
    }
    SampleTask* d_st = nullptr;
    float* d_cum = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_st, sizeof(SampleTask) * P, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_cum, 4 * (size_t)(off + 1), ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_st, st.data(), sizeof(SampleTask) * P, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mlp_sample_kernel, dim3((unsigned)P), dim3(256), 0, ctx->stream, d_st, P, d_cum, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(d_st, ctx->stream));
    HIP_TRY(hipFreeAsync(d_cum, ctx->stream));
    return 0;
}
}  // namespace cuacd

extern "C" {

int gpu_mlp_n_params(void) { return P_TOTAL; }

int gpu_mlp_create(gpu_ctx_t ctx, const float* params, int seed, gpu_mlp_t* out) {
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<float> w(P_TOTAL, 0.0f);
    if (params) {
        memcpy(w.data(), params, sizeof(float) * P_TOTAL);
    } else {
        // He-uniform weights on the used entries, zero biases and padding
        uint64_t s = 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)seed;
        auto rnd = [&]() {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            return (double)(s >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int rows_used, int rows, int cols_used, int cols, int fan_in) {
            const double a = std::sqrt(6.0 / fan_in);
            for (int r = 0; r < rows; ++r)
                for (int c = 0; c < cols; ++c)
                    w[off + r * cols + c] = (r < rows_used && c < cols_used) ? (float)((2.0 * rnd() - 1.0) * a) : 0.0f;
        };
        fill(P_W1, 4, KIN, HD, HD, 4);
        fill(P_W2, HD, HD, HD, HD, HD);
        fill(P_W3, HD, HD, HD, HD, HD);
        fill(P_H1, 2 * HD + 4, KH, HD, HD, 2 * HD + 4);
        fill(P_H2, HD, HD, HD, HD, HD);
        fill(P_H3, HD, HD, 1, NO, HD);
    }
    gpu_mlp* m = (gpu_mlp*)calloc(1, sizeof(gpu_mlp));
    REQUIRE(m, "out of host memory");
    const size_t fb = sizeof(float) * P_TOTAL;
    if (hipMalloc((void**)&m->w, fb) != hipSuccess || hipMalloc((void**)&m->wb, 2 * P_TOTAL) != hipSuccess ||
        hipMalloc((void**)&m->g, fb) != hipSuccess || hipMalloc((void**)&m->m, fb) != hipSuccess ||
        hipMalloc((void**)&m->v, fb) != hipSuccess) {
        gpu_mlp_destroy(ctx, m);
        set_error(ctx, "mlp: allocation failed");
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(m->w, w.data(), fb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(m->g, 0, fb, ctx->stream));
    HIP_TRY(hipMemsetAsync(m->m, 0, fb, ctx->stream));
    HIP_TRY(hipMemsetAsync(m->v, 0, fb, ctx->stream));
    hipLaunchKernelGGL(mlp_to_bf16, dim3((P_TOTAL + 255) / 256), dim3(256), 0, ctx->stream, m->w, m->wb, P_TOTAL);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = m;
    return 0;
}

void gpu_mlp_destroy(gpu_ctx_t ctx, gpu_mlp_t m) {
    if (!m) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    if (m->w) (void)hipFree(m->w);
    if (m->wb) (void)hipFree(m->wb);
    if (m->g) (void)hipFree(m->g);
    if (m->m) (void)hipFree(m->m);
    if (m->v) (void)hipFree(m->v);
    if (m->act) (void)hipFree(m->act);
    if (m->part) (void)hipFree(m->part);
    free(m);
}

int gpu_mlp_get_params(gpu_ctx_t ctx, gpu_mlp_t m, float* out) {
    REQUIRE(m, "null network");
    HIP_TRY(hipMemcpyAsync(out, m->w, sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gpu_mlp_forward(gpu_ctx_t ctx, gpu_mlp_t m, const float* points, int n_parts, const float* planes,
                    const int* part_of, int n_cand, float* out_scores) {
    REQUIRE(m && n_parts >= 0 && n_cand >= 0, "mlp_forward: bad arguments");
    if (n_cand == 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    Staged s;
    std::vector<int> coff;
    if (int rc = s.load(ctx, points, n_parts, planes, part_of, n_cand, nullptr, coff)) return stage_err(ctx, rc);
    float* d_out = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_out, 4ll * n_cand + 16, ctx->stream));
    int rc = mlp_forward_dev(ctx, m, s.pts, n_parts, s.planes, s.part, n_cand, d_out);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(out_scores, d_out, 4ll * n_cand, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    (void)hipFreeAsync(d_out, ctx->stream);
    return rc;
}

int gpu_mlp_train_step(gpu_ctx_t ctx, gpu_mlp_t m, const float* points, int n_parts, const float* planes,
                       const int* part_of, int n_cand, const float* targets, float lr, int apply,
                       float* out_loss, float* out_grads) {
    REQUIRE(m && n_parts > 0 && n_cand > 0, "mlp_train_step: bad arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    Staged s;
    std::vector<int> coff;
    if (int rc = s.load(ctx, points, n_parts, planes, part_of, n_cand, targets, coff)) return stage_err(ctx, rc);
    Act a;
    if (int rc = act_layout(ctx, m, n_cand, true, &a)) return rc;
    if (int rc = forward(ctx, m, s.pts, s.planes, s.part, n_cand, a)) return rc;
    if (int rc = backward(ctx, m, s.t, n_cand, a)) return rc;
    if (out_grads)
        HIP_TRY(hipMemcpyAsync(out_grads, m->g, sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost, ctx->stream));
    if (apply) {
        m->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(mlp_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, ctx->stream, m->w, m->g, m->m, m->v,
                           m->wb, P_TOTAL, lr, b1, b2, 1e-8f, c1, c2);
        HIP_TRY(hipGetLastError());
    }
    float loss = 0.0f;
    HIP_TRY(hipMemcpyAsync(&loss, a.loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (out_loss) *out_loss = loss;
    return 0;
}

int gpu_mlp_sample_parts(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
                         int n_parts, float* out_points) {
    REQUIRE(n_parts >= 0, "sample_parts: bad arguments");
    if (n_parts == 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t tv = (size_t)vert_off[n_parts], tt = (size_t)tri_off[n_parts];
    float* dv = nullptr;
    int* dt = nullptr;
    float* dout = nullptr;
    HIP_TRY(hipMallocAsync((void**)&dv, 12 * tv + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&dt, 12 * tt + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&dout, 12ull * NS * n_parts, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dv, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dt, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream));
    std::vector<MlpSamplePart> ps(n_parts);
    for (int p = 0; p < n_parts; ++p) {
        const int v0 = vert_off[p], v1 = vert_off[p + 1];
        REQUIRE(v1 >= v0 && tri_off[p + 1] >= tri_off[p], "sample_parts: bad offsets");
        ps[p].v = dv + 3ll * v0;
        ps[p].t = dt + 3ll * tri_off[p];
        ps[p].nv = v1 - v0;
        ps[p].nt = tri_off[p + 1] - tri_off[p];
        for (int j = 0; j < 3; ++j) { ps[p].lo[j] = 1e30f; ps[p].hi[j] = -1e30f; }
        for (int i = v0; i < v1; ++i)
            for (int j = 0; j < 3; ++j) {
                ps[p].lo[j] = std::min(ps[p].lo[j], verts[3ll * i + j]);
                ps[p].hi[j] = std::max(ps[p].hi[j], verts[3ll * i + j]);
            }
        for (int i = tri_off[p]; i < tri_off[p + 1]; ++i)
            for (int j = 0; j < 3; ++j)
                REQUIRE(tris[3ll * i + j] >= 0 && tris[3ll * i + j] < v1 - v0, "sample_parts: triangle index out of range");
    }
    int rc = mlp_sample_dev(ctx, ps.data(), n_parts, dout);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(out_points, dout, 12ull * NS * n_parts, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    (void)hipFreeAsync(dv, ctx->stream);
    (void)hipFreeAsync(dt, ctx->stream);
    (void)hipFreeAsync(dout, ctx->stream);
    return rc;
}

// GEMM through the C-ABI (tests of the MFMA tile layout): C [M x N] FP32 =
// A . B for BF16 host matrices with element strides.
int gpu_mlp_gemm_test(gpu_ctx_t ctx, int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                      long long a_len, const unsigned short* B, long long sbk, long long sbn, long long b_len,
                      float* C) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(M > 0 && N > 0 && K > 0 && a_len > 0 && b_len > 0, "gemm_test: bad sizes");
    REQUIRE((M - 1) * sam + (K - 1) * sak < a_len && (K - 1) * sbk + (N - 1) * sbn < b_len,
            "gemm_test: strides exceed the operands");
    bf16_t *dA = nullptr, *dB = nullptr;
    float* dC = nullptr;
    HIP_TRY(hipMallocAsync((void**)&dA, 2 * a_len, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&dB, 2 * b_len, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&dC, 4ll * M * N, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dA, A, 2 * a_len, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dB, B, 2 * b_len, hipMemcpyHostToDevice, ctx->stream));
    gpu_mlp tmp{};  // holds the split-K workspace of this call
    int rc = gemm(ctx, &tmp, M, N, K, dA, sam, sak, dB, sbk, sbn, nullptr, 0, nullptr, 0, dC, N, nullptr, 0);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(C, dC, 4ll * M * N, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    (void)hipFreeAsync(dA, ctx->stream);
    (void)hipFreeAsync(dB, ctx->stream);
    (void)hipFreeAsync(dC, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    if (tmp.part) (void)hipFree(tmp.part);
    return rc;
}

}  // extern "C"
