// fitsne_knn_mfma.hpp -- k nearest neighbours for large N on FP32 MFMA, with the exact FP64 result.
//
// The FP64 kernel of fitsne_knn.hpp (the VP tree's distance, vptree.h:97-107, rounded the same way)
// costs N^2 D FP64 operations with scattered candidate reads; at N in the millions it takes hours.  Here
// the same neighbours come from three passes:
//  1. threshold: the squared distance of every query to a sample of the points (every 16th) by FP32
//     MFMA (v_mfma_f32_16x16x4_f32, |x|^2 + |y|^2 - 2 x.y); tau_q = the 16th smallest, so about 256 of
//     all N points lie below it;
//  2. candidates: the same distances to all N points; every point with an approximate squared distance
//     below tau_q is appended to the query's list (at most kKmCap);
//  3. exact: the listed candidates' FP64 distances in the reference's order and rounding, sorted by
//     (distance, index) like the FP64 kernel, the first K kept.
// The approximate distance differs from the exact one by at most eps_q = 2 (Dp + 8) 2^-24 (|x_q|^2 +
// max_c |x_c|^2) (FP32 rounding of the inputs, the dot product over Dp terms, the norms and the final
// sum).  If the exact K-th distance is below tau_q - eps_q, no point left out can be among the K nearest,
// so the result is the FP64 kernel's bit for bit.  A query whose list is short, overflows or fails that
// margin is recomputed by the FP64 kernel.
//
// MFMA operand layout (16x16x4, FP32, one block): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// the result D[4 (l / 16) + v][l % 16], v < 4, is in lane l's accumulator v.  A holds 16 queries, B 16
// candidates, the inner dimension 4 coordinates; X is stored as float4 chunks of 4 coordinates,
// XT4[c][n] = x_n[4c .. 4c + 3] (zero padded), so that a wavefront's operand loads are contiguous.
#pragma once

#include <cfloat>
#include <climits>

#include "fitsne_common.hpp"
#include "fitsne_knn.hpp"

namespace fitsne {

constexpr int kKmT = 4;          // query tiles of 16 per wavefront (64 queries)
constexpr int kKmWaves = 4;      // wavefronts per workgroup
constexpr int kKmQ = kKmT * 16;  // queries per wavefront
constexpr int kKmStride = 16;    // the threshold sample: every 16th point
constexpr int kKmSampleK = 16;   // tau = the 16th smallest sample distance
constexpr int kKmCap = 512;      // candidates kept per query
constexpr int kKmSbuf = 64;      // threshold pass: per-query LDS buffer (16 kept + appends)

typedef float f32x4k __attribute__((ext_vector_type(4)));

// XT4[c][n] (c < nch) = FP32 coordinates 4c .. 4c + 3 of point n (zeros past D); nrm[n] = FP32 of the FP64
// squared norm (summed in coordinate order), nrm64[n] the FP64 value.  Every stride-th point also goes
// to the sample arrays (XS4[c][s], nrs[s]).
__global__ __launch_bounds__(256) void k_knnm_prep(const double *__restrict__ X, int N, int D, int nch,
                                                   float4 *__restrict__ XT4, float *__restrict__ nrm,
                                                   double *__restrict__ nrm64, int NS, float4 *__restrict__ XS4,
                                                   float *__restrict__ nrs) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const double *x = X + (long)n * D;
    double s = 0;
    const bool samp = n % kKmStride == 0 && n / kKmStride < NS;
    for (int c = 0; c < nch; c++) {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int d = 4 * c + k;
            const double xv = d < D ? x[d] : 0.0;
            s += xv * xv;
            v[k] = (float)xv;
        }
        const float4 f = make_float4(v[0], v[1], v[2], v[3]);
        XT4[(long)c * N + n] = f;
        if (samp) XS4[(long)c * NS + n / kKmStride] = f;
    }
    nrm[n] = (float)s;
    nrm64[n] = s;
    if (samp) nrs[n / kKmStride] = (float)s;
}

// max of v[0 .. n) into out[0] (one workgroup of 1024).
__global__ __launch_bounds__(1024) void k_knnm_max(const double *__restrict__ v, long n, double *__restrict__ out) {
    __shared__ double sh[1024 / kWave];
    double m = 0;
    for (long i = threadIdx.x; i < n; i += 1024) m = fmax(m, v[i]);
    m = wave_max(m);
    if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 1024 / kWave; i++) m = fmax(m, sh[i]);
        out[0] = m;
    }
}

// One wavefront's sweep over candidates [0, Nc) of XC4 (norms nc) for its queries q0 .. q0 + 63 of XQ4
// (Nq points, norms nq): for each tile of 16 candidates, visit(t, v, j, d2) on every lane with the
// approximate squared distance d2 between query q0 + 16 t + 4 (l / 16) + v and candidate j (j = Nc for
// the padding lanes of the last tile).  Queries past Nq read as point Nq - 1 (the caller ignores them).
template <int NCH, typename F>
__device__ __forceinline__ void knnm_sweep(const float *__restrict__ XQ4, int Nq, const float *__restrict__ nqv,
                                           int q0, const float *__restrict__ XC4, int Nc,
                                           const float *__restrict__ ncv, F &&visit) {
    const int l = threadIdx.x % kWave, r = l & 15, g = l >> 4;
    float a[kKmT][NCH];
    float nq[kKmT][4];
#pragma unroll
    for (int t = 0; t < kKmT; t++) {
        const int qa = min(q0 + 16 * t + r, Nq - 1);
#pragma unroll
        for (int c = 0; c < NCH; c++) a[t][c] = XQ4[((long)c * Nq + qa) * 4 + g];
#pragma unroll
        for (int v = 0; v < 4; v++) nq[t][v] = nqv[min(q0 + 16 * t + 4 * g + v, Nq - 1)];
    }
    for (int j0 = 0; j0 < Nc; j0 += 16) {
        const int j = j0 + r;
        const int jc = min(j, Nc - 1);
        float b[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) b[c] = XC4[((long)c * Nc + jc) * 4 + g];
        const float ncj = ncv[jc];
        f32x4k acc[kKmT];
#pragma unroll
        for (int t = 0; t < kKmT; t++) acc[t] = f32x4k{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int t = 0; t < kKmT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][c], b[c], acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < kKmT; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) visit(t, v, j < Nc ? j : Nc, (nq[t][v] + ncj) - 2.f * acc[t][v]);
    }
}

// Sort the 64 values of a wavefront (one per lane) ascending; returns this lane's value.
__device__ __forceinline__ float wave_sort64(float x) {
    const int l = threadIdx.x % kWave;
    for (int k = 2; k <= 64; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const float y = __shfl_xor(x, j, kWave);
            const bool lower = (l & j) == 0, up = (l & k) == 0;
            x = (lower == up) ? fminf(x, y) : fmaxf(x, y);
        }
    return x;
}

// Pass 1: tau[q] = the kKmSampleK-th smallest approximate squared distance from q to the sample points
// other than q itself (+inf if the sample has fewer).  Workgroup: kKmWaves wavefronts of 64 queries.
template <int NCH>
__global__ __launch_bounds__(kKmWaves * kWave) void k_knnm_tau(const float *__restrict__ XT4, int N,
                                                               const float *__restrict__ nrm,
                                                               const float *__restrict__ XS4, int NS,
                                                               const float *__restrict__ nrs,
                                                               float *__restrict__ tau) {
    __shared__ float sbuf[kKmWaves][kKmQ][kKmSbuf];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave, g = l >> 4;
    const int q0 = (blockIdx.x * kKmWaves + w) * kKmQ;
    if (q0 >= N) return;  // whole wavefront (no workgroup barriers below)
    float (*buf)[kKmSbuf] = sbuf[w];
    int cnt[kKmT][4];
    float th[kKmT][4];
#pragma unroll
    for (int t = 0; t < kKmT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            cnt[t][v] = 0;
            th[t][v] = INFINITY;
        }
    // sorts query qq's buffer (cnt entries), keeps the smallest kKmSampleK, returns the new threshold
    auto compact = [&](int qq, int n) -> float {
        const float x = l < n ? buf[qq][l] : INFINITY;
        const float s = wave_sort64(x);
        if (l < kKmSampleK) buf[qq][l] = s;
        return __shfl(s, kKmSampleK - 1, kWave);
    };
    knnm_sweep<NCH>(XT4, N, nrm, q0, XS4, NS, nrs, [&](int t, int v, int js, float d2) {
        const int q = q0 + 16 * t + 4 * g + v;
        const bool take = js < NS && js * kKmStride != q && d2 < th[t][v];
        const unsigned long long m = __ballot(take);
        const unsigned seg = (unsigned)(m >> (16 * g)) & 0xffffu;
        const int below = __popc(seg & ((1u << (l & 15)) - 1u));
        const int qq = 16 * t + 4 * g + v;
        if (take) buf[qq][cnt[t][v] + below] = d2;
        cnt[t][v] += __popc(seg);
        // a tile adds at most 16 per query: keep room for the next one
        const bool full = cnt[t][v] > kKmSbuf - 16;
        unsigned long long fm = __ballot(full && (l & 15) == 0);
        while (fm) {
            const int lead = __ffsll((long long)fm) - 1;
            const int gg = lead >> 4;
            const int n = __shfl(cnt[t][v], lead, kWave);
            const float nt = compact(16 * t + 4 * gg + v, n);
            if (g == gg) {
                cnt[t][v] = kKmSampleK;
                th[t][v] = nt;
            }
            fm &= fm - 1;
        }
    });
    // final threshold of every query: the kKmSampleK-th smallest of what its buffer holds
    for (int qq = 0; qq < kKmQ; qq++) {
        const int t = qq / 16, gg = (qq % 16) / 4, v = qq % 4;
        int n = 0;
#pragma unroll
        for (int tt = 0; tt < kKmT; tt++)
#pragma unroll
            for (int vv = 0; vv < 4; vv++)
                if (tt == t && vv == v) n = cnt[tt][vv];
        n = __shfl(n, 16 * gg, kWave);
        const float x = l < n ? buf[qq][l] : INFINITY;
        const float s = wave_sort64(x);
        const float tq = __shfl(s, kKmSampleK - 1, kWave);
        const int q = q0 + qq;
        if (l == 0 && q < N) tau[q] = tq;
    }
}

// Pass 2: every point j != q with approximate squared distance below tau[q] into cand[q][0 .. kKmCap);
// count[q] = how many there were (may exceed kKmCap: the query is then recomputed exactly).
template <int NCH>
__global__ __launch_bounds__(kKmWaves * kWave) void k_knnm_collect(const float *__restrict__ XT4, int N,
                                                                   const float *__restrict__ nrm,
                                                                   const float *__restrict__ tau,
                                                                   int *__restrict__ cand, int *__restrict__ count) {
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave, g = l >> 4;
    const int q0 = (blockIdx.x * kKmWaves + w) * kKmQ;
    if (q0 >= N) return;
    int cnt[kKmT][4];
    float th[kKmT][4];
#pragma unroll
    for (int t = 0; t < kKmT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            cnt[t][v] = 0;
            th[t][v] = tau[min(q0 + 16 * t + 4 * g + v, N - 1)];
        }
    knnm_sweep<NCH>(XT4, N, nrm, q0, XT4, N, nrm, [&](int t, int v, int j, float d2) {
        const int q = q0 + 16 * t + 4 * g + v;
        const bool take = j < N && j != q && q < N && d2 < th[t][v];
        const unsigned long long m = __ballot(take);
        const unsigned seg = (unsigned)(m >> (16 * g)) & 0xffffu;
        const int pos = cnt[t][v] + __popc(seg & ((1u << (l & 15)) - 1u));
        if (take && pos < kKmCap) cand[(long)q * kKmCap + pos] = j;
        cnt[t][v] += __popc(seg);
    });
#pragma unroll
    for (int t = 0; t < kKmT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int q = q0 + 16 * t + 4 * g + v;
            if ((l & 15) == 0 && q < N) count[q] = cnt[t][v];
        }
}

// Pass 3: one workgroup of one wavefront per query.  The listed candidates' FP64 squared distances
// (coordinate order, no fused multiply-add, as the FP64 kernel), sorted by (distance, index); the first
// K go to out_idx / out_dist (square root correctly rounded).  A query whose list is short or overflowed,
// or whose K-th distance does not clear tau - eps, is appended to redo[] instead.
__global__ __launch_bounds__(kWave) void k_knnm_exact(const double *__restrict__ X, int N, int D, int K,
                                                      const int *__restrict__ cand, const int *__restrict__ count,
                                                      const float *__restrict__ tau, const double *__restrict__ nrm64,
                                                      const double *__restrict__ nmax, double epsu,
                                                      int *__restrict__ out_idx, double *__restrict__ out_dist,
                                                      int *__restrict__ redo, int *__restrict__ nredo) {
    __shared__ double kk[kKmCap];
    __shared__ int ii[kKmCap];
    const int l = threadIdx.x;
    const int q = blockIdx.x;
    const int n = count[q];
    if (n < K || n > kKmCap) {
        if (l == 0) redo[atomicAdd(nredo, 1)] = q;
        return;
    }
    int B = 64;
    while (B < n) B <<= 1;
    const double *xq = X + (long)q * D;
    for (int i = l; i < B; i += kWave) {
        if (i < n) {
            const int c = cand[(long)q * kKmCap + i];
            const double *xc = X + (long)c * D;
            double acc = 0.0;
            for (int d = 0; d < D; d++) {
#pragma clang fp contract(off)
                const double diff = xq[d] - xc[d];
                acc += diff * diff;
            }
            kk[i] = acc;
            ii[i] = c;
        } else {
            kk[i] = DBL_MAX;
            ii[i] = INT_MAX;
        }
    }
    __syncthreads();
    knn_sort(kk, ii, B);  // by (distance, index), as the FP64 kernel's buffers
    const double dk = kk[K - 1];
    const double eps = epsu * (nrm64[q] + nmax[0]);
    if (!(dk < (double)tau[q] - eps)) {
        if (l == 0) redo[atomicAdd(nredo, 1)] = q;
        return;
    }
    for (int m = l; m < K; m += kWave) {
        out_idx[(long)q * K + m] = ii[m];
        out_dist[(long)q * K + m] = __dsqrt_rn(kk[m]);
    }
}

}  // namespace fitsne
