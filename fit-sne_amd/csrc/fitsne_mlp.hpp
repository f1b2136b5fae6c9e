// fitsne_mlp.hpp -- the learned warm start of the perplexity calibration: a small MLP on BF16 MFMA.
//
// For each point the model maps its K (<= 96) sorted kNN distances to a starting beta for the β
// bisection of distances2similarities (tsne.cpp:1394-1469), which then runs unchanged and remains the
// judge of the entropy error.  Network (FP32 master parameters, BF16 operands, FP32 accumulation):
//   x  = bf16((log(max(d^2, 1e-30)) - mu) / sd)          [96]  (features past K are 0)
//   h1 = relu(W1^T x + b1)                               [64]  W1 [96][64]
//   h2 = relu(W2^T bf16(h1) + b2)                        [64]  W2 [64][64]
//   y  = w3 . h2 + b3  = predicted log(beta)                   (FP32)
// Training minimises mean (y - log beta_target)^2 with Adam.
//
// One wavefront handles 16 samples.  Forward products use v_mfma_f32_16x16x32_bf16 (lane l holds
// A[l%16][8(l/16)+i] and B[8(l/16)+i][l%16], i < 8; D[4(l/16)+v][l%16], v < 4).  The weight gradients
// reduce over samples, so they use v_mfma_f32_16x16x16_bf16 with the 16 samples of the wavefront as
// the inner dimension (A[l%16][4(l/16)+i], B[4(l/16)+i][l%16], i < 4).  Gradient partials stay in the
// wavefront's registers over all its tiles, are combined over the 4 wavefronts in a fixed order in LDS,
// and written per workgroup; a second kernel sums the workgroup partials in order and applies Adam.
// Every step is deterministic.
#pragma once

#include "fitsne_common.hpp"

namespace fitsne {

constexpr int kMlpIn = 96;   // padded input width
constexpr int kMlpHid = 64;  // hidden width
constexpr int kMlpThreads = 256;
constexpr int kMlpWaves = kMlpThreads / kWave;
// FP32 parameter vector: mu[96] sd[96] W1[96][64] b1[64] W2[64][64] b2[64] w3[64] b3
constexpr int kMlpOffMu = 0, kMlpOffSd = kMlpIn, kMlpOffW1 = 2 * kMlpIn;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHid;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid;
constexpr int kMlpParams = kMlpOffB3 + 1;
// trainable part: W1 .. b3 (mu and sd are fixed from the training features)
constexpr int kMlpTrain0 = kMlpOffW1;
constexpr int kMlpTrainN = kMlpParams - kMlpOffW1;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4m __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    // round to nearest even (inputs are finite)
    const unsigned int u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }

// Standardised BF16 features from distances (row-major [n][K], FP64) into X [n][96].
__global__ __launch_bounds__(kThreads) void k_mlp_features(const double *__restrict__ dist, int n, int K,
                                                           const float *__restrict__ prm,
                                                           unsigned short *__restrict__ X) {
    const long e = (long)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (long)n * kMlpIn) return;
    const long s = e / kMlpIn;
    const int k = (int)(e - s * kMlpIn);
    float v = 0.f;
    if (k < K) {
        const double d = dist[s * K + k];
        const float f = (float)log(fmax(d * d, 1e-30));
        v = (f - prm[kMlpOffMu + k]) / prm[kMlpOffSd + k];
    }
    X[e] = bf16_bits(v);
}

// BF16 operand copies of the weights: W1T [64][96] (forward B of layer 1), W2T [64][64] (forward B of
// layer 2: W2T[j][a] = W2[a][j]) and W2 [64][64] (backward B of dh1 = dz2 W2^T).
__global__ __launch_bounds__(kThreads) void k_mlp_pack(const float *__restrict__ prm, unsigned short *__restrict__ w1t,
                                                       unsigned short *__restrict__ w2t,
                                                       unsigned short *__restrict__ w2) {
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e < kMlpHid * kMlpIn) {
        const int j = e / kMlpIn, k = e - j * kMlpIn;
        w1t[e] = bf16_bits(prm[kMlpOffW1 + k * kMlpHid + j]);
    }
    if (e < kMlpHid * kMlpHid) {
        const int j = e / kMlpHid, a = e - j * kMlpHid;
        w2t[e] = bf16_bits(prm[kMlpOffW2 + a * kMlpHid + j]);
        w2[e] = bf16_bits(prm[kMlpOffW2 + e]);
    }
}

struct MlpLds {
    unsigned short w1t[kMlpHid][kMlpIn];
    unsigned short w2t[kMlpHid][kMlpHid];
    unsigned short w2[kMlpHid][kMlpHid];
    float b1[kMlpHid], b2[kMlpHid], w3[kMlpHid];
    unsigned short x[kMlpWaves][16][kMlpIn];
    unsigned short h1[kMlpWaves][16][kMlpHid];
    unsigned short dz[kMlpWaves][16][kMlpHid];
};

__device__ void mlp_load_weights(MlpLds &L, const unsigned short *__restrict__ w1t, const unsigned short *__restrict__ w2t,
                                 const unsigned short *__restrict__ w2, const float *__restrict__ prm, bool need_w2) {
    for (int e = threadIdx.x; e < kMlpHid * kMlpIn; e += kMlpThreads) (&L.w1t[0][0])[e] = w1t[e];
    for (int e = threadIdx.x; e < kMlpHid * kMlpHid; e += kMlpThreads) {
        (&L.w2t[0][0])[e] = w2t[e];
        if (need_w2) (&L.w2[0][0])[e] = w2[e];
    }
    for (int e = threadIdx.x; e < kMlpHid; e += kMlpThreads) {
        L.b1[e] = prm[kMlpOffB1 + e];
        L.b2[e] = prm[kMlpOffB2 + e];
        L.w3[e] = prm[kMlpOffW3 + e];
    }
}

__device__ __forceinline__ bf16x8 lds_bf16x8(const unsigned short *p) {
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = __builtin_bit_cast(__bf16, p[i]);
    return v;
}

// Forward pass of one wavefront's 16 samples (rows s0 .. s0+15 of X, rows past n are zero).  Leaves
// z1 (pre-activation of layer 1) and z2 in D layout, bf16(h1) in L.h1[w], and y for samples
// 4(l/16)+v in yv[v] (identical across the 16 lanes of a quarter).
__device__ void mlp_forward_tile(MlpLds &L, int w, const unsigned short *__restrict__ X, long s0, long n, float b3,
                                 f32x4m (&z1)[4], f32x4m (&z2)[4], float (&yv)[4]) {
    const int l = threadIdx.x % kWave, q = l / 16, r = l % 16;
    // stage the X tile (16 x 96 bf16) in LDS
    for (int e = l; e < 16 * kMlpIn; e += kWave) {
        const int s = e / kMlpIn, k = e - s * kMlpIn;
        (&L.x[w][0][0])[e] = (s0 + s < n) ? X[(s0 + s) * kMlpIn + k] : (unsigned short)0;
    }
    __syncthreads();
    // layer 1: 3 K-chunks of 32, 4 output blocks of 16
#pragma unroll
    for (int nb = 0; nb < 4; nb++) z1[nb] = f32x4m{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < kMlpIn / 32; kc++) {
        const bf16x8 a = lds_bf16x8(&L.x[w][r][kc * 32 + 8 * q]);
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            const bf16x8 b = lds_bf16x8(&L.w1t[nb * 16 + r][kc * 32 + 8 * q]);
            z1[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, z1[nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            z1[nb][v] += L.b1[nb * 16 + r];
            L.h1[w][4 * q + v][nb * 16 + r] = bf16_bits(fmaxf(z1[nb][v], 0.f));
        }
    __syncthreads();
    // layer 2
#pragma unroll
    for (int nb = 0; nb < 4; nb++) z2[nb] = f32x4m{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < kMlpHid / 32; kc++) {
        const bf16x8 a = lds_bf16x8(&L.h1[w][r][kc * 32 + 8 * q]);
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            const bf16x8 b = lds_bf16x8(&L.w2t[nb * 16 + r][kc * 32 + 8 * q]);
            z2[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, z2[nb], 0, 0, 0);
        }
    }
    // layer 3 (FP32): partial dot over this lane's 4 units, then over the 16 lanes of the quarter
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float acc = 0.f;
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            z2[nb][v] += L.b2[nb * 16 + r];
            acc += fmaxf(z2[nb][v], 0.f) * L.w3[nb * 16 + r];
        }
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) acc += __shfl_xor(acc, o, kWave);
        yv[v] = acc + b3;
    }
}

// Inference: beta0[s] = exp(y_s).  Grid: one workgroup per 64 samples.
__global__ __launch_bounds__(kMlpThreads) void k_mlp_predict(const unsigned short *__restrict__ X, long n,
                                                             const unsigned short *__restrict__ w1t,
                                                             const unsigned short *__restrict__ w2t,
                                                             const float *__restrict__ prm,
                                                             double *__restrict__ beta0) {
    __shared__ MlpLds L;
    mlp_load_weights(L, w1t, w2t, nullptr, prm, false);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const long s0 = ((long)blockIdx.x * kMlpWaves + w) * 16;
    f32x4m z1[4], z2[4];
    float yv[4];
    __syncthreads();
    mlp_forward_tile(L, w, X, s0, n, prm[kMlpOffB3], z1, z2, yv);
    if (l % 16 == 0) {
        const int q = l / 16;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const long s = s0 + 4 * q + v;
            if (s < n) beta0[s] = exp((double)yv[v]);
        }
    }
}

// Gradient of the mean squared error over samples [0, n) (loss normaliser inv_n = 1/n_total) for one
// workgroup: tiles t = blockIdx.x, blockIdx.x + gridDim.x, ... of 64 samples.  Writes the workgroup's
// partial gradient (trainable layout, kMlpTrainN floats) and partial loss sum to gpart / lpart.
__global__ __launch_bounds__(kMlpThreads) void k_mlp_grad(const unsigned short *__restrict__ X,
                                                          const float *__restrict__ target, long n, float inv_n,
                                                          const unsigned short *__restrict__ w1t,
                                                          const unsigned short *__restrict__ w2t,
                                                          const unsigned short *__restrict__ w2,
                                                          const float *__restrict__ prm, float *__restrict__ gpart,
                                                          double *__restrict__ lpart) {
    __shared__ MlpLds L;
    __shared__ float red[kMlpWaves][kWave];
    mlp_load_weights(L, w1t, w2t, w2, prm, true);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave, q = l / 16, r = l % 16;
    const float b3 = prm[kMlpOffB3];
    f32x4m gW1[6][4], gW2[4][4];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) gW1[i][j] = f32x4m{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) gW2[i][j] = f32x4m{0.f, 0.f, 0.f, 0.f};
    float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f}, gw3[4] = {0.f, 0.f, 0.f, 0.f};
    float gb3 = 0.f;
    double loss = 0.0;
    __syncthreads();
    const long ntiles = (n + 63) / 64;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long s0 = t * 64 + w * 16;
        f32x4m z1[4], z2[4];
        float yv[4];
        mlp_forward_tile(L, w, X, s0, n, b3, z1, z2, yv);
        float dy[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const long s = s0 + 4 * q + v;
            const float e = s < n ? yv[v] - target[s] : 0.f;
            dy[v] = 2.f * e * inv_n;
            if (r == 0) loss += (double)e * e;
        }
        // output layer and dz2 = dy w3 [z2 > 0]
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float h2 = fmaxf(z2[nb][v], 0.f);
                gw3[nb] += dy[v] * h2;
                const float dz = z2[nb][v] > 0.f ? dy[v] * L.w3[nb * 16 + r] : 0.f;
                gb2[nb] += dz;
                L.dz[w][4 * q + v][nb * 16 + r] = bf16_bits(dz);
            }
        if (r == 0)
#pragma unroll
            for (int v = 0; v < 4; v++) gb3 += dy[v];
        __syncthreads();
        // gW2[a][j] += sum_s h1[s][a] dz2[s][j]: A = h1^T (a x s), B = dz2 (s x j)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            s16x4 a;
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = (short)L.h1[w][4 * q + i][mb * 16 + r];
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                s16x4 b;
#pragma unroll
                for (int i = 0; i < 4; i++) b[i] = (short)L.dz[w][4 * q + i][nb * 16 + r];
                gW2[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, gW2[mb][nb], 0, 0, 0);
            }
        }
        // dh1 = dz2 W2^T (s x a): A = dz2 (s x j), B[j][a] = W2[a][j]; then dz1 = dh1 [z1 > 0]
        f32x4m dh1[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) dh1[nb] = f32x4m{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < kMlpHid / 32; kc++) {
            const bf16x8 a = lds_bf16x8(&L.dz[w][r][kc * 32 + 8 * q]);
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                const bf16x8 b = lds_bf16x8(&L.w2[nb * 16 + r][kc * 32 + 8 * q]);
                dh1[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, dh1[nb], 0, 0, 0);
            }
        }
        __syncthreads();  // all lanes have read dz2 before it is overwritten with dz1
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float dz = z1[nb][v] > 0.f ? dh1[nb][v] : 0.f;
                gb1[nb] += dz;
                L.dz[w][4 * q + v][nb * 16 + r] = bf16_bits(dz);
            }
        __syncthreads();
        // gW1[k][j] += sum_s x[s][k] dz1[s][j]: A = x^T (k x s), B = dz1 (s x j)
#pragma unroll
        for (int mb = 0; mb < 6; mb++) {
            s16x4 a;
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = (short)L.x[w][4 * q + i][mb * 16 + r];
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                s16x4 b;
#pragma unroll
                for (int i = 0; i < 4; i++) b[i] = (short)L.dz[w][4 * q + i][nb * 16 + r];
                gW1[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, gW1[mb][nb], 0, 0, 0);
            }
        }
        __syncthreads();  // the tile's LDS rows are reused by the next tile
    }
    // combine the 4 wavefronts in a fixed order, one trainable parameter per (lane, slot)
    float *out = gpart + (long)blockIdx.x * kMlpTrainN;
    auto combine = [&](float v, int idx) {
        // idx: trainable index; the 4 wavefronts' values are summed wave 0 + 1 + 2 + 3
        red[w][l] = v;
        __syncthreads();
        if (w == 0) out[idx] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
        __syncthreads();
    };
    // W1: D layout of block (mb, nb): element [k = mb*16 + 4q + v][j = nb*16 + r]
#pragma unroll
    for (int mb = 0; mb < 6; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                combine(gW1[mb][nb][v], (mb * 16 + 4 * q + v) * kMlpHid + nb * 16 + r);
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                combine(gW2[mb][nb][v], (kMlpOffW2 - kMlpTrain0) + (mb * 16 + 4 * q + v) * kMlpHid + nb * 16 + r);
    // biases and w3: sum over the 4 quarters of the wavefront first (fixed order), then over wavefronts
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        float b1v = gb1[nb], b2v = gb2[nb], w3v = gw3[nb];
        b1v += __shfl_xor(b1v, 16, kWave);
        b1v += __shfl_xor(b1v, 32, kWave);
        b2v += __shfl_xor(b2v, 16, kWave);
        b2v += __shfl_xor(b2v, 32, kWave);
        w3v += __shfl_xor(w3v, 16, kWave);
        w3v += __shfl_xor(w3v, 32, kWave);
        // lanes r (q = 0) now hold unit nb*16 + r
        red[w][l] = b1v;
        __syncthreads();
        if (w == 0 && q == 0) out[(kMlpOffB1 - kMlpTrain0) + nb * 16 + r] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
        __syncthreads();
        red[w][l] = b2v;
        __syncthreads();
        if (w == 0 && q == 0) out[(kMlpOffB2 - kMlpTrain0) + nb * 16 + r] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
        __syncthreads();
        red[w][l] = w3v;
        __syncthreads();
        if (w == 0 && q == 0) out[(kMlpOffW3 - kMlpTrain0) + nb * 16 + r] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
        __syncthreads();
    }
    // b3 and the loss: lanes r == 0 of each quarter hold partials
    red[w][l] = gb3;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int ww = 0; ww < kMlpWaves; ww++)
            for (int qq = 0; qq < 4; qq++) s += red[ww][16 * qq];
        out[kMlpOffB3 - kMlpTrain0] = s;
    }
    __syncthreads();
    __shared__ double lred[kMlpWaves][kWave];
    lred[w][l] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int ww = 0; ww < kMlpWaves; ww++)
            for (int qq = 0; qq < 4; qq++) s += lred[ww][16 * qq];
        lpart[blockIdx.x] = s;
    }
}

// Sum of the workgroup partials in order -> grad (trainable layout); optional Adam step on prm.
__global__ __launch_bounds__(kThreads) void k_mlp_reduce_adam(const float *__restrict__ gpart, int nparts,
                                                              float *__restrict__ grad, float *__restrict__ prm,
                                                              float *__restrict__ m1, float *__restrict__ m2,
                                                              int do_step, float lr, float c1, float c2) {
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= kMlpTrainN) return;
    float g = 0.f;
    for (int p = 0; p < nparts; p++) g += gpart[(long)p * kMlpTrainN + e];
    grad[e] = g;
    if (do_step) {
        // Adam (beta1 0.9, beta2 0.999, eps 1e-8); c1, c2 = 1 - beta^t bias corrections
        const float a = 0.9f * m1[e] + 0.1f * g;
        const float b = 0.999f * m2[e] + 0.001f * g * g;
        m1[e] = a;
        m2[e] = b;
        prm[kMlpTrain0 + e] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
    }
}

}  // namespace fitsne
