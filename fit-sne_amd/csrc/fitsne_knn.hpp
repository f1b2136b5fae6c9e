// fitsne_knn.hpp -- exact k nearest neighbours in FP64 for the affinity build (gfx950, wave64).
//
// TSNE::run finds each point's K nearest other points with a VP tree (tsne.cpp:1643-1726), whose
// distance is sqrt(sum_d (x_d - y_d)^2) accumulated in order d = 0..D-1 without fused multiply-adds
// (vptree.h:97-107).  Here one wavefront serves kKnnQ query points: every lane takes one candidate per
// round, accumulates its squared distances to the kKnnQ queries in the same order and rounding
// (contraction to fused multiply-adds switched off; the root by the correctly rounded __dsqrt_rn),
// and candidates below the query's current K-th distance are appended to an LDS buffer of B slots
// per query.  When a buffer is nearly full it is sorted by (distance, index)
// with a bitonic network and cut back to K, which tightens the threshold.  The output is sorted by
// (distance, index), so equal distances are broken by the smaller index and the result is
// deterministic.  X is read transposed (XT[d][n]) so that the lanes' candidate loads coalesce.
#pragma once

#include <cfloat>
#include <climits>

#include "fitsne_common.hpp"

namespace fitsne {

constexpr int kKnnQ = 4;

// XT[d][n] = X[n][d], 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose_f64(const double *__restrict__ X, long N, int D,
                                                       double *__restrict__ XT) {
    __shared__ double tile[32][33];
    const long n0 = (long)blockIdx.x * 32;
    const int d0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const long n = n0 + r;
        const int d = d0 + tx;
        tile[r][tx] = (n < N && d < D) ? X[n * D + d] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int d = d0 + r;
        const long n = n0 + tx;
        if (n < N && d < D) XT[(long)d * N + n] = tile[tx][r];
    }
}

// Sort key[0..B), idx[0..B) ascending by (key, idx); B a power of two; one wavefront.
__device__ void knn_sort(double *key, int *idx, int B) {
    const int lane = threadIdx.x;
    for (int k = 2; k <= B; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < B; i += kWave) {
                const int l = i ^ j;
                if (l > i) {
                    const double a = key[i], b = key[l];
                    const int ia = idx[i], ib = idx[l];
                    const bool gt = a > b || (a == b && ia > ib);
                    if (gt == ((i & k) == 0)) {
                        key[i] = b;
                        key[l] = a;
                        idx[i] = ib;
                        idx[l] = ia;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Pad a buffer's tail to B with (+inf, INT_MAX), sort it, keep the first K.
__device__ void knn_compact(double *key, int *idx, int cnt, int B) {
    for (int i = cnt + (int)threadIdx.x; i < B; i += kWave) {
        key[i] = DBL_MAX;
        idx[i] = INT_MAX;
    }
    __syncthreads();
    knn_sort(key, idx, B);
}

// One workgroup = one wavefront = kKnnQ queries: points blockIdx.x * kKnnQ + j, or, with a query list,
// qlist[blockIdx.x * kKnnQ + j] for list positions below nq.  Dynamic LDS: kKnnQ * B * (8 + 4) bytes,
// where B = power of two >= max(2K, K + 64, 128) so that a round of 64 appends always fits after a cut.
__global__ __launch_bounds__(kWave) void k_knn_f64(const double *__restrict__ X, const double *__restrict__ XT,
                                                   int N, int D, int K, int B, int *__restrict__ out_idx,
                                                   double *__restrict__ out_dist, const int *__restrict__ qlist,
                                                   int nq) {
    extern __shared__ double knn_smem[];
    double *bk = knn_smem;
    int *bi = reinterpret_cast<int *>(knn_smem + (long)kKnnQ * B);
    const int lane = threadIdx.x;
    int qid[kKnnQ];  // the block's queries (-1: none)
#pragma unroll
    for (int j = 0; j < kKnnQ; j++) {
        const int s = blockIdx.x * kKnnQ + j;
        qid[j] = qlist ? (s < nq ? qlist[s] : -1) : (s < N ? s : -1);
    }
    int cnt[kKnnQ];
    double thr[kKnnQ];
#pragma unroll
    for (int j = 0; j < kKnnQ; j++) {
        cnt[j] = 0;
        thr[j] = DBL_MAX;
    }
    for (int c0 = 0; c0 < N; c0 += kWave) {
        const int c = c0 + lane;
        double acc[kKnnQ];
#pragma unroll
        for (int j = 0; j < kKnnQ; j++) acc[j] = 0.0;
        if (c < N) {
            // separate multiply and add, as the host's loop (no fused multiply-add)
#pragma clang fp contract(off)
            for (int d = 0; d < D; d++) {
                const double xc = XT[(long)d * N + c];
#pragma unroll
                for (int j = 0; j < kKnnQ; j++) {
                    const int qi = qid[j] >= 0 ? qid[j] : 0;
                    const double diff = X[(long)qi * D + d] - xc;
                    acc[j] += diff * diff;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kKnnQ; j++) {
            const int qi = qid[j];
            const bool take = c < N && qi >= 0 && c != qi && acc[j] < thr[j];
            const unsigned long long m = __ballot(take);
            if (m) {
                const int pos = cnt[j] + __popcll(m & ((1ull << lane) - 1ull));
                if (take) {
                    bk[j * B + pos] = acc[j];
                    bi[j * B + pos] = c;
                }
                cnt[j] += __popcll(m);
                if (cnt[j] > B - kWave) {
                    __syncthreads();
                    knn_compact(bk + j * B, bi + j * B, cnt[j], B);
                    cnt[j] = K;
                    thr[j] = bk[j * B + K - 1];
                    __syncthreads();
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kKnnQ; j++) {
        const int qi = qid[j];
        knn_compact(bk + j * B, bi + j * B, cnt[j], B);
        if (qi >= 0)
            for (int m = lane; m < K; m += kWave) {
                out_idx[(long)qi * K + m] = bi[j * B + m];
                out_dist[(long)qi * K + m] = __dsqrt_rn(bk[j * B + m]);  // correctly rounded, as the host's sqrt
            }
        __syncthreads();
    }
}

}  // namespace fitsne
