// Learned PBMPM warm start: a 16 -> 64 -> 64 -> 12 MLP (GELU) with BF16 weights and FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x16_bf16).
//
// Per particle it maps the state at the start of a substep
//   x = [singular values of F_e, largest first (3), det F_p (1), D (9, column-major),
//        dprev (3): the previous substep's last per-iteration change of dx]                  (16)
// to the change that k_ref PBMPM iterations make,
//   y = [D_converged - D (9, column-major), dx_converged - dx - g dt^2 (3)]                 (12, pad 16)
// (g dt^2: what the iterations' gravity shares add to dx). The slow part of the PBMPM iteration is a
// smooth, body-wide mode whose per-iteration change of dx varies little from one substep to the
// next, so dprev carries most of where the iterations will take dx; a prediction of D alone does not
// move the residual. Features and targets are standardised with per-column mean and standard
// deviation of the logged training set.
//
// MFMA 16x16x16 bf16 operand layout (lane l of a wave64):
//   A[row = l % 16][k = 4 (l / 16) + j], B[k = 4 (l / 16) + j][col = l % 16],
//   D[row = 4 (l / 16) + i][col = l % 16],  i, j = 0..3.
// A wave processes 16 samples at a time. Training transposes the activations through LDS between
// layers (the backward pass reuses them there); inference computes the transposed products and
// keeps them in registers (mlp_infer16).
#pragma once
#include "mpm_kernels.h"

namespace mpm {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// parameter block (FP32 master copy), offsets in floats
enum : int {
    MLP_IN = 16, MLP_H = 64, MLP_OUT = 16, MLP_NIN = 16, MLP_NOUT = 12,
    MLP_W1 = 0,                       // [16][64]   row = input feature
    MLP_B1 = MLP_W1 + 16 * 64,        // [64]
    MLP_W2 = MLP_B1 + 64,             // [64][64]   row = layer-1 unit
    MLP_B2 = MLP_W2 + 64 * 64,        // [64]
    MLP_W3 = MLP_B2 + 64,             // [64][16]   row = layer-2 unit, columns 12..15 zero
    MLP_B3 = MLP_W3 + 64 * 16,        // [16]
    MLP_NPARAM = MLP_B3 + 16          // 6288
};

MPM_DEV short to_bf16(float x) {      // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
MPM_DEV float from_bf16(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// GELU (tanh form) and its derivative through the logistic function: 0.5 (1 + tanh u) = s(2u),
// s(a) = 1 / (1 + exp(-a)), u = sqrt(2/pi) (z + 0.044715 z^3), so GELU(z) = z s(2u) and
// GELU'(z) = s + 2 z s (1 - s) u'. exp(-2u) = exp2(z (K0 + K1 z^2)) with K0 = -2 sqrt(2/pi) log2(e),
// K1 = 0.044715 K0, on the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32, 1 ulp each): the
// network's arithmetic is BF16-limited, and 128 of these per sample are most of the warm start's
// VALU work. No cancellation for negative z (1 + tanh u is never formed), and exp2 -> inf for very
// negative z gives s = 0.
MPM_DEV float gelu_s(float z) {
    const float a = z * __builtin_fmaf(-0.10294324f, z * z, -2.3022082f);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a));
}
MPM_DEV float gelu(float z) { return z * gelu_s(z); }
MPM_DEV float gelu_grad(float z) {
    const float s = gelu_s(z);
    return s + ((2.0f * z) * (s * (1.0f - s))) * (0.7978845608f * (1.0f + 0.134145f * (z * z)));
}

struct MlpBufs {
    const float* params;      // FP32 master [MLP_NPARAM]
    const short* wbf;         // BF16 copies: W1 [16][64], W2 [64][64], W3 [64][16] (same layouts)
    float* grads;             // FP32 [MLP_NPARAM]
    const float* norm;        // [4][16]: feature mean, feature std, target mean, target std
};

// LDS images of the weights in MFMA B-operand order, 8 bytes per lane per tile (training's
// Z = X W; inference's Z^T = W^T X^T reads the same images as A operands).
//   w1b[n][l]      = W1[4(l/16)+j][16n + l%16]                     (layer 1, B = W1)
//   w2b[kk][n][l]  = W2[16kk + 4(l/16)+j][16n + l%16]              (layer 2, B = W2)
//   w3b[kk][l]     = W3[16kk + 4(l/16)+j][l%16]                    (layer 3, B = W3)
// The backward images (training only, MlpLdsBack):
//   w3t[n][l]      = W3[16n + l%16][4(l/16)+j]                     (backward dH2 = dY W3^T)
//   w2t[kk][n][l]  = W2[16n + l%16][16kk + 4(l/16)+j]              (backward dH1 = dZ2 W2^T)
// isd_x = 1 / sd_x: features are standardised as (x - mu_x) * isd_x. 16-byte aligned: inference
// reads the biases four at a time (b1, b2, b3 start at multiples of 16 bytes).
struct alignas(16) MlpLds {
    bf4 w1b[4][64];
    bf4 w2b[4][4][64];
    bf4 w3b[4][64];
    float b1[64], b2[64], b3[16];
    float mu_x[16], isd_x[16], mu_y[16], sd_y[16];
};
static_assert(offsetof(MlpLds, b1) % 16 == 0 && offsetof(MlpLds, b2) % 16 == 0 && offsetof(MlpLds, b3) % 16 == 0,
              "bias rows are read as 16-byte vectors");
struct MlpLdsBack {
    bf4 w3t[4][64];
    bf4 w2t[4][4][64];
};

MPM_DEV void mlp_stage_weights(MlpLds& L, MlpLdsBack* T, const MlpBufs& M) {
    const short* W1 = M.wbf;
    const short* W2 = M.wbf + 16 * 64;
    const short* W3 = M.wbf + 16 * 64 + 64 * 64;
    for (int e = threadIdx.x; e < 64 * 4; e += blockDim.x) {
        const int n = e / 64, l = e % 64;
        bf4 a, c, d;
        for (int j = 0; j < 4; ++j) {
            a[j] = W1[(4 * (l / 16) + j) * 64 + 16 * n + l % 16];
            c[j] = W3[(16 * n + 4 * (l / 16) + j) * 16 + l % 16];
            d[j] = W3[(16 * n + l % 16) * 16 + 4 * (l / 16) + j];
        }
        L.w1b[n][l] = a; L.w3b[n][l] = c;
        if (T) T->w3t[n][l] = d;
    }
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
        const int kk = e / 256, n = (e / 64) % 4, l = e % 64;
        bf4 a, b;
        for (int j = 0; j < 4; ++j) {
            a[j] = W2[(16 * kk + 4 * (l / 16) + j) * 64 + 16 * n + l % 16];
            b[j] = W2[(16 * n + l % 16) * 64 + 16 * kk + 4 * (l / 16) + j];
        }
        L.w2b[kk][n][l] = a;
        if (T) T->w2t[kk][n][l] = b;
    }
    for (int e = threadIdx.x; e < 64; e += blockDim.x) { L.b1[e] = M.params[MLP_B1 + e]; L.b2[e] = M.params[MLP_B2 + e]; }
    for (int e = threadIdx.x; e < 16; e += blockDim.x) {
        L.b3[e] = M.params[MLP_B3 + e];
        L.mu_x[e] = M.norm[e]; L.isd_x[e] = 1.0f / M.norm[16 + e]; L.mu_y[e] = M.norm[32 + e]; L.sd_y[e] = M.norm[48 + e];
    }
}

#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x16bf16_1k((a), (b), (c), 0, 0, 0)

// Forward pass of 16 samples whose normalised features are xs[s][0..15] (bf16) in LDS. Leaves
// z1, z2 (pre-activations, output layout), h1s / h2s (bf16 activations [s][64] in LDS) and y.
struct MlpAct {
    f4 z1[4], z2[4], y;
};
MPM_DEV void mlp_forward16(const MlpLds& L, const short* xs, short* h1s, short* h2s, MlpAct& A) {
    const int l = lane_id();
    const int r = l % 16, q = l / 16;
    bf4 xa;
    for (int j = 0; j < 4; ++j) xa[j] = xs[r * 16 + 4 * q + j];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        c = MFMA(xa, L.w1b[n][l], c);
        const float b = L.b1[16 * n + r];
        for (int i = 0; i < 4; ++i) { c[i] += b; h1s[(4 * q + i) * 64 + 16 * n + r] = to_bf16(gelu(c[i])); }
        A.z1[n] = c;
    }
    wave_sync();
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            bf4 ha;
            for (int j = 0; j < 4; ++j) ha[j] = h1s[r * 64 + 16 * kk + 4 * q + j];
            c = MFMA(ha, L.w2b[kk][n][l], c);
        }
        const float b = L.b2[16 * n + r];
        for (int i = 0; i < 4; ++i) c[i] += b;
        A.z2[n] = c;
    }
    wave_sync();
#pragma unroll
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < 4; ++i) h2s[(4 * q + i) * 64 + 16 * n + r] = to_bf16(gelu(A.z2[n][i]));
    wave_sync();
    f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        bf4 ha;
        for (int j = 0; j < 4; ++j) ha[j] = h2s[r * 64 + 16 * kk + 4 * q + j];
        c = MFMA(ha, L.w3b[kk][l], c);
    }
    const float b = L.b3[r];
    for (int i = 0; i < 4; ++i) c[i] += b;
    A.y = c;
}

// Per-wave LDS scratch of the 16-sample forward / backward.
struct MlpWave {
    short xs[16 * 16];
    short h1s[16 * 64];
    short h2s[16 * 64];
    short ts[16 * 64];        // transposes of dY / dZ2
};

// One training minibatch: each wave runs `tiles` tiles of 16 samples drawn with replacement
// (sample = hash1(seed ^ global tile / lane) mod n) and accumulates MSE gradients (of the
// standardised targets, mean over the minibatch) into M.grads; loss_sum gets sum of squared errors.
struct MlpTrainArgs {
    const float* feats;   // [n][16] raw features
    const float* targs;   // [n][16] raw targets
    uint64_t n;
    uint32_t seed;
    uint32_t tiles;       // tiles of 16 samples per wave
    float inv_batch;      // 1 / (samples per minibatch)
    float* loss_sum;
};
__global__ __launch_bounds__(256) void k_mlp_train(MlpBufs M, MlpTrainArgs T) {
    __shared__ MlpLds L;
    __shared__ MlpLdsBack LT;
    __shared__ MlpWave Wv[4];
    __shared__ float G[MLP_NPARAM];
    mlp_stage_weights(L, &LT, M);
    for (int e = threadIdx.x; e < MLP_NPARAM; e += blockDim.x) G[e] = 0.0f;
    __syncthreads();
    const int w = threadIdx.x >> 6, l = lane_id(), r = l % 16, q = l / 16;
    MlpWave& S = Wv[w];
    f4 g1[4], g2[4][4], g3[4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f, loss = 0.0f;
    for (int a = 0; a < 4; ++a) { g1[a] = f4{0, 0, 0, 0}; g3[a] = f4{0, 0, 0, 0}; for (int b = 0; b < 4; ++b) g2[a][b] = f4{0, 0, 0, 0}; }
    const uint32_t wave_global = blockIdx.x * 4u + (uint32_t)w;
    for (uint32_t t = 0; t < T.tiles; ++t) {
        const uint32_t tile = wave_global * T.tiles + t;
        // sample of row r: one 64-bit index per row, drawn from two 32-bit hashes
        const uint32_t h0 = hash1(T.seed ^ hash1(tile * 16u + (uint32_t)r));
        const uint32_t h1 = hash1(h0 ^ 0x9E3779B9u);
        const uint64_t s = (((uint64_t)h1 << 32) | h0) % T.n;
        // features: lane (r, q) loads 4 and normalises them
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * q + j;
            const float x = f < MLP_NIN ? (T.feats[s * 16 + f] - L.mu_x[f]) * L.isd_x[f] : 0.0f;
            S.xs[r * 16 + f] = to_bf16(x);
        }
        wave_sync();
        MlpAct A;
        mlp_forward16(L, S.xs, S.h1s, S.h2s, A);
        // dY in output layout: rows 4q+i (samples), column r (output); standardised targets
        f4 dy;
        for (int i = 0; i < 4; ++i) {
            const uint32_t hh0 = hash1(T.seed ^ hash1(tile * 16u + (uint32_t)(4 * q + i)));
            const uint32_t hh1 = hash1(hh0 ^ 0x9E3779B9u);
            const uint64_t si = (((uint64_t)hh1 << 32) | hh0) % T.n;
            float d = 0.0f;
            if (r < MLP_NOUT) {
                const float tt = (T.targs[si * 16 + r] - L.mu_y[r]) / L.sd_y[r];
                const float e = A.y[i] - tt;
                loss += e * e;
                d = (2.0f * e) * T.inv_batch;
            }
            dy[i] = d;
        }
        // dW3 += H2^T dY ; db3 += sum_s dY
        bf4 dyb;
        for (int i = 0; i < 4; ++i) { dyb[i] = to_bf16(dy[i]); gb3 += dy[i]; S.ts[(4 * q + i) * 16 + r] = dyb[i]; }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            bf4 ha;
            for (int j = 0; j < 4; ++j) ha[j] = S.h2s[(4 * q + j) * 64 + 16 * m + r];
            g3[m] = MFMA(ha, dyb, g3[m]);
        }
        wave_sync();
        // dH2 = dY W3^T, dZ2 = dH2 * gelu'(Z2)
        bf4 dya;
        for (int j = 0; j < 4; ++j) dya[j] = S.ts[r * 16 + 4 * q + j];
        f4 dz2[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            c = MFMA(dya, LT.w3t[n][l], c);
            for (int i = 0; i < 4; ++i) c[i] = c[i] * gelu_grad(A.z2[n][i]);
            dz2[n] = c;
        }
        wave_sync();
        // dW2 += H1^T dZ2 ; db2 ; stage dZ2 [s][64] for the next product
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            bf4 db;
            for (int i = 0; i < 4; ++i) { db[i] = to_bf16(dz2[n][i]); gb2[n] += dz2[n][i]; S.ts[(4 * q + i) * 64 + 16 * n + r] = db[i]; }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf4 ha;
                for (int j = 0; j < 4; ++j) ha[j] = S.h1s[(4 * q + j) * 64 + 16 * m + r];
                g2[m][n] = MFMA(ha, db, g2[m][n]);
            }
        }
        wave_sync();
        // dH1 = dZ2 W2^T, dZ1 = dH1 * gelu'(Z1); dW1 += X^T dZ1 ; db1
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                bf4 da;
                for (int j = 0; j < 4; ++j) da[j] = S.ts[r * 64 + 16 * kk + 4 * q + j];
                c = MFMA(da, LT.w2t[kk][n][l], c);
            }
            bf4 db;
            for (int i = 0; i < 4; ++i) { const float dz = c[i] * gelu_grad(A.z1[n][i]); db[i] = to_bf16(dz); gb1[n] += dz; }
            bf4 xa;
            for (int j = 0; j < 4; ++j) xa[j] = S.xs[(4 * q + j) * 16 + r];
            g1[n] = MFMA(xa, db, g1[n]);
        }
        wave_sync();
    }
    // reduce the wave partials into LDS, then one global atomic per parameter per workgroup
    for (int m = 0; m < 4; ++m) for (int i = 0; i < 4; ++i) {
        atomicAdd(&G[MLP_W1 + (4 * q + i) * 64 + 16 * m + r], g1[m][i]);            // rows: features
        atomicAdd(&G[MLP_W3 + (16 * m + 4 * q + i) * 16 + r], g3[m][i]);            // rows: layer-2 units
        for (int n = 0; n < 4; ++n) atomicAdd(&G[MLP_W2 + (16 * m + 4 * q + i) * 64 + 16 * n + r], g2[m][n][i]);
    }
    for (int n = 0; n < 4; ++n) { atomicAdd(&G[MLP_B1 + 16 * n + r], gb1[n]); atomicAdd(&G[MLP_B2 + 16 * n + r], gb2[n]); }
    atomicAdd(&G[MLP_B3 + r], gb3);
    loss = wave_sum_f(loss);
    if (l == 0) atomicAdd(T.loss_sum, loss);
    __syncthreads();
    for (int e = threadIdx.x; e < MLP_NPARAM; e += blockDim.x) if (G[e] != 0.0f) atomicAdd(&M.grads[e], G[e]);
}

// Adam step on the FP32 master parameters; refreshes the BF16 copies and clears the gradients.
__global__ void k_mlp_adam(float* params, float* grads, float* m1, float* m2, short* wbf, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MLP_NPARAM) return;
    const float g = grads[e];
    const float a = b1 * m1[e] + (1.0f - b1) * g;
    const float v = b2 * m2[e] + (1.0f - b2) * (g * g);
    m1[e] = a; m2[e] = v;
    const float p = params[e] - lr * (a / c1) / (__builtin_sqrtf(v / c2) + eps);
    params[e] = p;
    grads[e] = 0.0f;
    if (e < MLP_B1) wbf[e] = to_bf16(p);
    else if (e >= MLP_W2 && e < MLP_B2) wbf[16 * 64 + (e - MLP_W2)] = to_bf16(p);
    else if (e >= MLP_W3 && e < MLP_B3) wbf[16 * 64 + 64 * 64 + (e - MLP_W3)] = to_bf16(p);
}
__global__ void k_mlp_refresh_bf16(const float* params, short* wbf) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MLP_NPARAM) return;
    const float p = params[e];
    if (e < MLP_B1) wbf[e] = to_bf16(p);
    else if (e >= MLP_W2 && e < MLP_B2) wbf[16 * 64 + (e - MLP_W2)] = to_bf16(p);
    else if (e >= MLP_W3 && e < MLP_B3) wbf[16 * 64 + 64 * 64 + (e - MLP_W3)] = to_bf16(p);
}

// Gradient of the mean squared error over exactly n samples (given order, no sampling): test hook
// for checking the MFMA forward / backward against a host computation. One wave per 16 samples.
__global__ __launch_bounds__(256) void k_mlp_grad_fixed(MlpBufs M, const float* feats, const float* targs, uint32_t n,
                                                        float* out_y, float* loss_sum) {
    __shared__ MlpLds L;
    __shared__ MlpLdsBack LT;
    __shared__ MlpWave Wv[4];
    mlp_stage_weights(L, &LT, M);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = lane_id(), r = l % 16, q = l / 16;
    MlpWave& S = Wv[w];
    const uint32_t tile = blockIdx.x * 4u + (uint32_t)w;
    if (tile * 16u >= n) return;                 // n is a multiple of 16 (checked on the host)
    const float inv_n = 1.0f / (float)n;
    for (int j = 0; j < 4; ++j) {
        const int f = 4 * q + j;
        const float x = f < MLP_NIN ? (feats[(size_t)(tile * 16 + r) * 16 + f] - L.mu_x[f]) * L.isd_x[f] : 0.0f;
        S.xs[r * 16 + f] = to_bf16(x);
    }
    wave_sync();
    MlpAct A;
    mlp_forward16(L, S.xs, S.h1s, S.h2s, A);
    f4 dy;
    float loss = 0.0f;
    for (int i = 0; i < 4; ++i) {
        const uint32_t si = tile * 16 + 4 * q + i;
        out_y[(size_t)si * 16 + r] = A.y[i];
        float d = 0.0f;
        if (r < MLP_NOUT) {
            const float e = A.y[i] - (targs[(size_t)si * 16 + r] - L.mu_y[r]) / L.sd_y[r];
            loss += e * e;
            d = (2.0f * e) * inv_n;
        }
        dy[i] = d;
    }
    bf4 dyb;
    for (int i = 0; i < 4; ++i) { dyb[i] = to_bf16(dy[i]); S.ts[(4 * q + i) * 16 + r] = dyb[i]; atomicAdd(&M.grads[MLP_B3 + r], dy[i]); }
    for (int m = 0; m < 4; ++m) {
        bf4 ha;
        for (int j = 0; j < 4; ++j) ha[j] = S.h2s[(4 * q + j) * 64 + 16 * m + r];
        f4 g = {0.f, 0.f, 0.f, 0.f};
        g = MFMA(ha, dyb, g);
        for (int i = 0; i < 4; ++i) atomicAdd(&M.grads[MLP_W3 + (16 * m + 4 * q + i) * 16 + r], g[i]);
    }
    wave_sync();
    bf4 dya;
    for (int j = 0; j < 4; ++j) dya[j] = S.ts[r * 16 + 4 * q + j];
    f4 dz2[4];
    for (int n2 = 0; n2 < 4; ++n2) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        c = MFMA(dya, LT.w3t[n2][l], c);
        for (int i = 0; i < 4; ++i) c[i] = c[i] * gelu_grad(A.z2[n2][i]);
        dz2[n2] = c;
    }
    wave_sync();
    for (int n2 = 0; n2 < 4; ++n2) {
        bf4 db;
        for (int i = 0; i < 4; ++i) { db[i] = to_bf16(dz2[n2][i]); S.ts[(4 * q + i) * 64 + 16 * n2 + r] = db[i]; atomicAdd(&M.grads[MLP_B2 + 16 * n2 + r], dz2[n2][i]); }
        for (int m = 0; m < 4; ++m) {
            bf4 ha;
            for (int j = 0; j < 4; ++j) ha[j] = S.h1s[(4 * q + j) * 64 + 16 * m + r];
            f4 g = {0.f, 0.f, 0.f, 0.f};
            g = MFMA(ha, db, g);
            for (int i = 0; i < 4; ++i) atomicAdd(&M.grads[MLP_W2 + (16 * m + 4 * q + i) * 64 + 16 * n2 + r], g[i]);
        }
    }
    wave_sync();
    for (int n2 = 0; n2 < 4; ++n2) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < 4; ++kk) {
            bf4 da;
            for (int j = 0; j < 4; ++j) da[j] = S.ts[r * 64 + 16 * kk + 4 * q + j];
            c = MFMA(da, LT.w2t[kk][n2][l], c);
        }
        bf4 db;
        for (int i = 0; i < 4; ++i) { const float dz = c[i] * gelu_grad(A.z1[n2][i]); db[i] = to_bf16(dz); atomicAdd(&M.grads[MLP_B1 + 16 * n2 + r], dz); }
        bf4 xa;
        for (int j = 0; j < 4; ++j) xa[j] = S.xs[(4 * q + j) * 16 + r];
        f4 g = {0.f, 0.f, 0.f, 0.f};
        g = MFMA(xa, db, g);
        for (int i = 0; i < 4; ++i) atomicAdd(&M.grads[MLP_W1 + (4 * q + i) * 64 + 16 * n2 + r], g[i]);
    }
    loss = wave_sum_f(loss);
    if (l == 0) atomicAdd(loss_sum, loss);
}

// Features of a particle at the start of a substep (see the header comment).
// The singular values are those of the exact 6-sweep Jacobi SVD (svd3, the SVD plasticity option's).
// An eigenvalue-only Jacobi on the hardware rcp / sqrt / rsqrt costs a tenth as much and agrees to
// ~1e-7 in a float32 emulation, but warm starts trained on its features ran unstably (end-of-run
// error 3-25x that of the same network on these features, and varying from one training to the next).
MPM_DEV void mlp_features(const m3& Fe, float Jp, const m3& D, v3 dprev, float f[16]) {
    float a[9], v[9], b[9], sig[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { a[3 * c] = Fe.c[c].x; a[3 * c + 1] = Fe.c[c].y; a[3 * c + 2] = Fe.c[c].z; }
    svd3<float>(a, v, b, sig);
    float s0 = sig[0], s1 = sig[1], s2 = sig[2], t;        // largest first
    if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
    if (s1 < s2) { t = s1; s1 = s2; s2 = t; }
    if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
    f[0] = s0; f[1] = s1; f[2] = s2;
    f[3] = Jp;
    for (int c = 0; c < 3; ++c) { f[4 + 3 * c] = D.c[c].x; f[5 + 3 * c] = D.c[c].y; f[6 + 3 * c] = D.c[c].z; }
    f[13] = dprev.x; f[14] = dprev.y; f[15] = dprev.z;
}
MPM_DEV v3 mlp_dprev(const Bufs& B, uint32_t i) {
    return B.dprev ? ld3(B.dprev, B.stride, 0, i) : mk(0.0f, 0.0f, 0.0f);
}

// Logs the features of every particle (current order) into the dataset rows base + particle id,
// and the particle's dx into columns 12..14 of its target row (read back by k_mlp_log_targets).
__global__ __launch_bounds__(256) void k_mlp_log_features(Bufs B, uint32_t n, float* feats, float* targs, uint64_t base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sd = B.stride;
    float f[16];
    mlp_features(ldm(B.cur, sd, FE, i), ld(B.cur, sd, JP, i), ldm(B.cur, sd, DD, i), mlp_dprev(B, i), f);
    const uint64_t row = base + B.pid_cur[i];
    float4* o = reinterpret_cast<float4*>(feats + row * 16);
    o[0] = make_float4(f[0], f[1], f[2], f[3]); o[1] = make_float4(f[4], f[5], f[6], f[7]);
    o[2] = make_float4(f[8], f[9], f[10], f[11]); o[3] = make_float4(f[12], f[13], f[14], f[15]);
    const v3 dx = ld3(B.cur, sd, DXX, i);
    reinterpret_cast<float4*>(targs + row * 16)[3] = make_float4(dx.x, dx.y, dx.z, 0.0f);
}
// After the substep (state in `cur`): target = (D_converged - D_start, dx_converged - dx_start -
// g dt^2), D_start = features 4..12, dx_start = target columns 12..14 (logged with the features),
// dx_converged = the last G2P's dx (resid_dx, before integrate's walls and collider). A particle that
// no iteration reached this substep (dprev 0) gets a zero dx target.
__global__ __launch_bounds__(256) void k_mlp_log_targets(Bufs B, uint32_t n, const float* feats, float* targs, uint64_t base,
                                                         float gx, float gy, float gz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sd = B.stride;
    const uint64_t row = base + B.pid_cur[i];
    const m3 D = ldm(B.cur, sd, DD, i);
    float t[16];
    for (int c = 0; c < 3; ++c) {
        t[3 * c] = D.c[c].x - feats[row * 16 + 4 + 3 * c];
        t[3 * c + 1] = D.c[c].y - feats[row * 16 + 5 + 3 * c];
        t[3 * c + 2] = D.c[c].z - feats[row * 16 + 6 + 3 * c];
    }
    const v3 dp = mlp_dprev(B, i);
    const bool iterated = dp.x != 0.0f || dp.y != 0.0f || dp.z != 0.0f;
    const v3 dxc = ld3(B.resid_dx, sd, 0, i);
    const float* dx0 = targs + row * 16 + 12;
    t[9] = iterated ? (dxc.x - dx0[0]) - gx : 0.0f;
    t[10] = iterated ? (dxc.y - dx0[1]) - gy : 0.0f;
    t[11] = iterated ? (dxc.z - dx0[2]) - gz : 0.0f;
    for (int c = 12; c < 16; ++c) t[c] = 0.0f;
    float4* o = reinterpret_cast<float4*>(targs + row * 16);
    for (int c = 0; c < 4; ++c) o[c] = make_float4(t[4 * c], t[4 * c + 1], t[4 * c + 2], t[4 * c + 3]);
}

// Column moments of the dataset for standardisation, FP32 with compensated (Kahan) summation:
// each workgroup writes its 64 partial sums (x, x^2, t, t^2 per column, 16 columns each) to
// part[blockIdx.x * 64 ..]; the host adds the partials in double.
__global__ __launch_bounds__(256) void k_mlp_colstats(const float* feats, const float* targs, uint64_t n, float* part) {
    __shared__ float red[16][64];
    float acc[4] = {0, 0, 0, 0}, cmp[4] = {0, 0, 0, 0};
    const int c = threadIdx.x & 15;
    const uint64_t rows_per = (gridDim.x * (uint64_t)blockDim.x) / 16;
    for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 16; r < n; r += rows_per) {
        const float x = feats[r * 16 + c], t = targs[r * 16 + c];
        const float v[4] = {x, x * x, t, t * t};
        for (int k = 0; k < 4; ++k) {
            const float y = v[k] - cmp[k];
            const float s = acc[k] + y;
            cmp[k] = (s - acc[k]) - y;
            acc[k] = s;
        }
    }
    // 16 threads share each column: reduce them in LDS (rows of 16 lanes)
    const int row = threadIdx.x >> 4;
    for (int k = 0; k < 4; ++k) red[row][16 * k + c] = acc[k];
    __syncthreads();
    if (threadIdx.x < 64) {
        float s = 0.0f, comp = 0.0f;
        for (int r = 0; r < 16; ++r) {
            const float y = red[r][threadIdx.x] - comp;
            const float t = s + y;
            comp = (t - s) - y;
            s = t;
        }
        part[blockIdx.x * 64u + threadIdx.x] = s;
    }
}

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
// Four floats to bfloat16, round to nearest even (v_cvt_pk_bf16_f32), as the B operand of an MFMA.
MPM_DEV bf4 bf16x4(float a, float b, float c, float d) {
    const bf2v lo = __builtin_convertvector((f2v){a, b}, bf2v);
    const bf2v hi = __builtin_convertvector((f2v){c, d}, bf2v);
    const uint32_t u0 = __builtin_bit_cast(uint32_t, lo), u1 = __builtin_bit_cast(uint32_t, hi);
    bf4 o;
    o[0] = (short)(u0 & 0xFFFFu); o[1] = (short)(u0 >> 16); o[2] = (short)(u1 & 0xFFFFu); o[3] = (short)(u1 >> 16);
    return o;
}

// Inference forward of 16 samples with every product transposed, Z^T = W^T X^T: the weight images
// above are the A operands (A[row = unit][k] = W[k][unit] is the lane image w1b / w2b / w3b), the
// activations the B operands. The D layout of one layer's output block (lane l: units 4(l/16)+i of
// the block, sample l%16) is the B layout that the next layer reads (k = 4(l/16)+j, column l%16), so
// the activations stay in registers from layer to layer. xb: lane l holds the standardised
// features 4(l/16)..4(l/16)+3 of sample l%16. Returns outputs 4(l/16)..4(l/16)+3 of sample l%16.
MPM_DEV f4 mlp_infer16(const MlpLds& L, bf4 xb) {
    const int l = lane_id(), q = l / 16;
    bf4 h[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        c = MFMA(L.w1b[n][l], xb, c);
        const f4 b = *reinterpret_cast<const f4*>(&L.b1[16 * n + 4 * q]);
        h[n] = bf16x4(gelu(c[0] + b[0]), gelu(c[1] + b[1]), gelu(c[2] + b[2]), gelu(c[3] + b[3]));
    }
    bf4 h2[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) c = MFMA(L.w2b[kk][n][l], h[kk], c);
        const f4 b = *reinterpret_cast<const f4*>(&L.b2[16 * n + 4 * q]);
        h2[n] = bf16x4(gelu(c[0] + b[0]), gelu(c[1] + b[1]), gelu(c[2] + b[2]), gelu(c[3] + b[3]));
    }
    f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) c = MFMA(L.w3b[kk][l], h2[kk], c);
    const f4 b = *reinterpret_cast<const f4*>(&L.b3[4 * q]);
    return c + b;
}

// Warm start of a PBMPM substep, per particle in `cur`: D += the predicted D change and dx += the
// predicted dx change (de-standardised). A particle falls back to the plain method (prediction
// dropped) when any standardised output exceeds `limit` in magnitude or is not finite, or when no
// iteration reached it in the previous substep (dprev 0: no input to predict from); the count of
// fallbacks goes to counters[C_FALLBACK].
// Per wave and 64 particles (lane l owns particle base + l): the features into xs (row l, bf16),
// 4 tiles of 16 through mlp_infer16, the outputs into ys (row = particle, 12 floats: rows 48 B apart
// keep the 16-byte writes and reads free of bank conflicts), then each lane applies its own row.
struct alignas(16) MlpInferWave {  // ys rows are written 16 bytes at a time
    bf4 xs[64][4];            // [particle][features 4g..4g+3]
    float ys[64][MLP_NOUT];
};
__global__ __launch_bounds__(256, 4) void k_mlp_warmstart(MlpBufs M, Bufs B, uint32_t n, float limit, uint32_t* n_fallback) {
    __shared__ MlpLds L;
    __shared__ MlpInferWave Wv[4];
    __shared__ uint32_t n_fb;                 // this workgroup's fallbacks: one global atomic at the end
    mlp_stage_weights(L, nullptr, M);
    if (threadIdx.x == 0) n_fb = 0u;
    __syncthreads();
    const int w = threadIdx.x >> 6, l = lane_id(), r = l % 16, q = l / 16;
    MlpInferWave& S = Wv[w];
    const uint32_t sd = B.stride;
    for (uint32_t base = (blockIdx.x * 4u + (uint32_t)w) * 64u; base < n; base += gridDim.x * 256u) {
        const uint32_t i = base + (uint32_t)l;
        float f[16];
        m3 D = zero3();
        v3 dp = mk(0.0f, 0.0f, 0.0f);
        if (i < n) {
            D = ldm(B.cur, sd, DD, i);
            dp = mlp_dprev(B, i);
            mlp_features(ldm(B.cur, sd, FE, i), ld(B.cur, sd, JP, i), D, dp, f);
        } else {
            for (int c = 0; c < 16; ++c) f[c] = L.mu_x[c];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g)
            S.xs[l][g] = bf16x4((f[4 * g] - L.mu_x[4 * g]) * L.isd_x[4 * g], (f[4 * g + 1] - L.mu_x[4 * g + 1]) * L.isd_x[4 * g + 1],
                                (f[4 * g + 2] - L.mu_x[4 * g + 2]) * L.isd_x[4 * g + 2], (f[4 * g + 3] - L.mu_x[4 * g + 3]) * L.isd_x[4 * g + 3]);
        wave_sync();
        for (int t = 0; t < 4; ++t) {                     // 4 tiles of 16 particles
            const f4 y = mlp_infer16(L, S.xs[16 * t + r][q]);
            if (q < MLP_NOUT / 4) *reinterpret_cast<f4*>(&S.ys[16 * t + r][4 * q]) = y;
        }
        wave_sync();
        bool ok = true;
        if (i < n) {
            float dd[MLP_NOUT];
            for (int c = 0; c < MLP_NOUT; ++c) {
                const float ys_c = S.ys[l][c];
                if (!(ys_c <= limit && ys_c >= -limit)) ok = false;
                dd[c] = ys_c * L.sd_y[c] + L.mu_y[c];
            }
            if (dp.x == 0.0f && dp.y == 0.0f && dp.z == 0.0f) ok = false;
            if (ok) {
                for (int c = 0; c < 3; ++c) { D.c[c].x += dd[3 * c]; D.c[c].y += dd[3 * c + 1]; D.c[c].z += dd[3 * c + 2]; }
                stm(B.cur, sd, DD, i, D);
                st3(B.cur, sd, DXX, i, ld3(B.cur, sd, DXX, i) + mk(dd[9], dd[10], dd[11]));
            }
        }
        const uint64_t fb = __ballot(i < n && !ok);
        if (l == 0 && fb) atomicAdd(&n_fb, (uint32_t)__popcll(fb));
        wave_sync();
    }
    __syncthreads();
    if (threadIdx.x == 0 && n_fb) atomicAdd(n_fallback, n_fb);
}

}  // namespace mpm
