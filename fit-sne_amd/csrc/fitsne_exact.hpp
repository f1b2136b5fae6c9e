// fitsne_exact.hpp -- exact t-SNE in FP64 (theta = 0): dense affinities, O(N^2) gradient and KL
// (gfx950, wave64).
//
// The reference's exact mode keeps the whole N x N matrix P on the host (tsne.cpp:171-229):
//   * computeGaussianPerplexity, dense overload (tsne.cpp:1503-1530): squared distances
//     (computeSquaredEuclideanDistance, tsne.cpp:1832-1848), then each row through the beta bisection of
//     distances2similarities with the row's own entry set to DBL_MIN (tsne.cpp:1421, :1459);
//   * P[n][m] = P[m][n] = P[n][m] + P[m][n] for n < m, then P /= sum(P) (tsne.cpp:196-216);
//   * computeExactGradient (tsne.cpp:1232-1282) and evaluateError (tsne.cpp:1286-1326).
// Here P stays in HBM.  The gradient is one pass over P: a workgroup owns kExRows rows, its threads
// sweep the columns (coalesced row segments of P, Y[m] shared by the rows), and each thread accumulates
// per row the attractive sum  sum_m P q (y_n - y_m), the repulsive sum  sum_m q^df q (y_n - y_m)  and
// the row's share of Z = sum_{m != n} q^df; dC = att - rep / Z afterwards.  That is the reference's
// sum_m (P - q^df / Z) q (y_n - y_m) regrouped, so P is read once instead of once per pass.  All
// reductions have a fixed order: results are bitwise repeatable.
#pragma once

#include <cfloat>

#include "fitsne_kernels.hpp"

namespace fitsne {

constexpr int kExRows = 4;     // rows of P per workgroup in the gradient and KL sweeps
constexpr int kExMaxDims = 4;  // largest embedding dimension with a compiled exact kernel
constexpr int kSqRows = 8;     // rows per workgroup of the squared-distance matrix
constexpr int kSymTile = 32;   // tile edge of the symmetrisation

// DD[n][m] = sum_d (X[n][d] - X[m][d])^2, summed from 0.0 in order d = 0..D-1 without fused
// multiply-adds, as the reference's loop (tsne.cpp:1832-1848); (a - b)^2 == (b - a)^2 in IEEE
// arithmetic, so the lower triangle equals the reference's mirrored copy bit for bit.  XT[d][m] = X[m][d]
// keeps the column loads coalesced.  Grid: (ceil(N / kThreads), ceil(N / kSqRows)).
__global__ __launch_bounds__(kThreads) void k_sqdist_f64(const double *__restrict__ X, const double *__restrict__ XT,
                                                         int N, int D, double *__restrict__ DD) {
    const int m = blockIdx.x * kThreads + threadIdx.x;
    const int n0 = blockIdx.y * kSqRows;
    if (m >= N) return;
    double acc[kSqRows];
#pragma unroll
    for (int r = 0; r < kSqRows; r++) acc[r] = 0.0;
    {
#pragma clang fp contract(off)
        for (int d = 0; d < D; d++) {
            const double xm = XT[(long)d * N + m];
#pragma unroll
            for (int r = 0; r < kSqRows; r++) {
                const int n = min(n0 + r, N - 1);
                const double diff = X[(long)n * D + d] - xm;
                acc[r] += diff * diff;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kSqRows; r++) {
        const int n = n0 + r;
        if (n < N) DD[(long)n * N + m] = acc[r];
    }
}

// P[n][m] = P[m][n] = P[n][m] + P[m][n] for n < m, diagonal untouched (tsne.cpp:196-207).  Workgroup
// (I, J), J >= I, owns the tiles (I, J) and (J, I): both are read into LDS before either is written,
// and no other workgroup touches them.  Grid: (T, T) with T = ceil(N / kSymTile); blocks below the
// diagonal exit.  256 threads = 32 x 8.
__global__ __launch_bounds__(256) void k_exact_sym(double *__restrict__ P, int N) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    __shared__ double a[kSymTile][kSymTile + 1], b[kSymTile][kSymTile + 1];
    const int tx = threadIdx.x % kSymTile, ty = threadIdx.x / kSymTile;
    const long r0 = (long)I * kSymTile, c0 = (long)J * kSymTile;
    for (int r = ty; r < kSymTile; r += 256 / kSymTile) {
        const long n = r0 + r, m = c0 + tx;
        a[r][tx] = (n < N && m < N) ? P[n * N + m] : 0.0;  // a[r][c] = P[r0 + r][c0 + c]
        const long n2 = c0 + r, m2 = r0 + tx;
        b[r][tx] = (n2 < N && m2 < N) ? P[n2 * N + m2] : 0.0;  // b[r][c] = P[c0 + r][r0 + c]
    }
    __syncthreads();
    for (int r = ty; r < kSymTile; r += 256 / kSymTile) {
        // upper entry (r0 + r, c0 + tx): P[n][m] + P[m][n]
        const long n = r0 + r, m = c0 + tx;
        if (n < N && m < N && n < m) P[n * N + m] = a[r][tx] + b[tx][r];
        // lower entry (c0 + r, r0 + tx): the sum of its upper partner (r0 + tx, c0 + r)
        const long n2 = c0 + r, m2 = r0 + tx;
        if (n2 < N && m2 < N && m2 < n2) P[n2 * N + m2] = a[tx][r] + b[r][tx];
    }
}

// Block partial sums of v[0..n) over a fixed grid (grid-stride, fixed order).
__global__ __launch_bounds__(kThreads) void k_partial_sum(const double *__restrict__ v, long n,
                                                          double *__restrict__ part) {
    double s = 0.0;
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads) s += v[e];
    __shared__ double sh[kThreads / kWave];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// v[e] /= s[0] (the normalisation P /= sum_P, tsne.cpp:215).
__global__ __launch_bounds__(kThreads) void k_div_by(double *__restrict__ v, long n, const double *__restrict__ s) {
    const double d = s[0];
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads) v[e] /= d;
}

// q = 1 / (1 + d2 / df) with the reference's operations (tsne.cpp:1256); d2 / 1.0 == d2 exactly, and
// q^1 == q, so the df = 1 instantiation (DF1) leaves out the division by df and the power.
template <bool DF1>
__device__ __forceinline__ double exact_q(double d2, double df) {
    if constexpr (DF1) return 1.0 / (1.0 + d2);
    else return 1.0 / (1.0 + d2 / df);
}
template <bool DF1>
__device__ __forceinline__ double exact_qpow(double q, double df) {
    if constexpr (DF1) return q;
    else return pow(q, df);
}

// One sweep over rows n0 .. n0 + kExRows - 1 of the pair matrix.  GRAD: att[n], rep[n] (D each) and
// zrow[n] = sum_{m != n} q^df; otherwise zrow only (P unused).  Rows past N (last workgroup) are
// computed on row N - 1 and discarded.  The next column's Y and P values are loaded into a second
// register buffer while the current column is computed, so the loads of consecutive columns overlap
// (the P loads are unconditional: every index is in bounds, the self pair is masked afterwards).
// DF1 = (df == 1).
// Grid: ceil(N / kExRows) workgroups of kThreads.
template <int D, bool DF1, bool GRAD>
__global__ __launch_bounds__(kThreads) void k_exact_rows(const double *__restrict__ P, const double *__restrict__ Y,
                                                         int N, double df, double *__restrict__ att,
                                                         double *__restrict__ rep, double *__restrict__ zrow) {
    const int n0 = blockIdx.x * kExRows;
    double yn[kExRows][D], a[kExRows][D], b[kExRows][D], s[kExRows];
    const double *prow[kExRows];
#pragma unroll
    for (int r = 0; r < kExRows; r++) {
        const int n = min(n0 + r, N - 1);
        prow[r] = GRAD ? P + (long)n * N : nullptr;
#pragma unroll
        for (int d = 0; d < D; d++) {
            yn[r][d] = Y[(long)n * D + d];
            a[r][d] = 0.0;
            b[r][d] = 0.0;
        }
        s[r] = 0.0;
    }
    // two register buffers: the loads of column m + kThreads are in flight while column m is computed
    double yA[D], pA[kExRows], yB[D], pB[kExRows];
    // branch-free loads (a column past N reloads column N - 1 and is not used), so that the wait for
    // one buffer never waits for the other
    auto fetch = [&](int mm, double (&y)[D], double (&p)[kExRows]) {
        const int mc = min(mm, N - 1);
#pragma unroll
        for (int d = 0; d < D; d++) y[d] = Y[(long)mc * D + d];
        if constexpr (GRAD) {
#pragma unroll
            for (int r = 0; r < kExRows; r++) p[r] = prow[r][mc];
        }
    };
    auto column = [&](int mm, const double (&ym)[D], const double (&pv)[kExRows]) {
#pragma unroll
        for (int r = 0; r < kExRows; r++) {
            const bool self = mm == n0 + r;
            double diff[D], d2 = 0.0;
            {
#pragma clang fp contract(off)
#pragma unroll
                for (int d = 0; d < D; d++) {
                    diff[d] = yn[r][d] - ym[d];
                    d2 += diff[d] * diff[d];
                }
            }
            const double q = exact_q<DF1>(d2, df);
            const double qp = exact_qpow<DF1>(q, df);
            s[r] += self ? 0.0 : qp;
            if constexpr (GRAD) {
                const double pq = (self ? 0.0 : pv[r]) * q;
                const double qq = self ? 0.0 : qp * q;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    a[r][d] += diff[d] * pq;
                    b[r][d] += diff[d] * qq;
                }
            }
        }
    };
    int m = threadIdx.x;
    fetch(m, yA, pA);
    for (; m < N; m += 2 * kThreads) {
        fetch(m + kThreads, yB, pB);
        column(m, yA, pA);
        if (m + kThreads >= N) break;
        fetch(m + 2 * kThreads, yA, pA);
        column(m + kThreads, yB, pB);
    }
    constexpr int V = 2 * D + 1;
    __shared__ double red[kThreads / kWave][kExRows * V];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int r = 0; r < kExRows; r++) {
        if constexpr (GRAD) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double ta = wave_sum(a[r][d]);
                const double tb = wave_sum(b[r][d]);
                if (l == 0) {
                    red[w][r * V + d] = ta;
                    red[w][r * V + D + d] = tb;
                }
            }
        }
        const double ts = wave_sum(s[r]);
        if (l == 0) red[w][r * V + 2 * D] = ts;
    }
    __syncthreads();
    if (threadIdx.x < kExRows * V) {
        const int r = threadIdx.x / V, k = threadIdx.x % V;
        const int n = n0 + r;
        if (n < N && (GRAD || k == 2 * D)) {
            double t = 0.0;
            for (int i = 0; i < kThreads / kWave; i++) t += red[i][threadIdx.x];
            if (k == 2 * D) zrow[n] = t;
            else if (k < D) att[(long)n * D + k] = t;
            else rep[(long)n * D + k - D] = t;
        }
    }
}

// dC = att - rep / Z, Z = zbuf[0] (the regrouped tsne.cpp:1268-1279).
__global__ __launch_bounds__(kThreads) void k_exact_combine(const double *__restrict__ att,
                                                            const double *__restrict__ rep,
                                                            const double *__restrict__ zbuf, long n,
                                                            double *__restrict__ dC) {
    const double Z = zbuf[0];
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads)
        dC[e] = att[e] - rep[e] / Z;
}

// KL of evaluateError (tsne.cpp:1286-1326) over rows n0 .. n0 + kExRows - 1: Q = q^df / sum_Q off the
// diagonal, DBL_MIN / sum_Q on it, sum_Q = DBL_MIN + sum_{n != m} q^df (zsum[0] holds the sum);
// C += P log((P + FLT_MIN) / (Q + FLT_MIN)) over all N^2 entries.  Block partials into part.
template <int D, bool DF1>
__global__ __launch_bounds__(kThreads) void k_exact_kl(const double *__restrict__ P, const double *__restrict__ Y,
                                                       int N, double df, const double *__restrict__ zsum,
                                                       double *__restrict__ part) {
    const int n0 = blockIdx.x * kExRows;
    const double sQ = DBL_MIN + zsum[0];
    const double fm = (double)FLT_MIN;
    double yn[kExRows][D];
    long prow[kExRows];
#pragma unroll
    for (int r = 0; r < kExRows; r++) {
        const int n = min(n0 + r, N - 1);
        prow[r] = (long)n * N;
#pragma unroll
        for (int d = 0; d < D; d++) yn[r][d] = Y[(long)n * D + d];
    }
    double acc = 0.0;
    double yA[D], pA[kExRows], yB[D], pB[kExRows];
    auto fetch = [&](int mm, double (&y)[D], double (&p)[kExRows]) {
        const int mc = min(mm, N - 1);
#pragma unroll
        for (int d = 0; d < D; d++) y[d] = Y[(long)mc * D + d];
#pragma unroll
        for (int r = 0; r < kExRows; r++) p[r] = P[prow[r] + mc];
    };
    auto column = [&](int mm, const double (&ym)[D], const double (&pv)[kExRows]) {
#pragma unroll
        for (int r = 0; r < kExRows; r++) {
            double d2 = 0.0;
            {
#pragma clang fp contract(off)
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double t = yn[r][d] - ym[d];
                    d2 += t * t;
                }
            }
            const double Q = (mm == n0 + r ? DBL_MIN : exact_qpow<DF1>(exact_q<DF1>(d2, df), df)) / sQ;
            const double p = pv[r];
            const double c = p * log((p + fm) / (Q + fm));
            acc += n0 + r < N ? c : 0.0;
        }
    };
    int m = threadIdx.x;
    fetch(m, yA, pA);
    for (; m < N; m += 2 * kThreads) {
        fetch(m + kThreads, yB, pB);
        column(m, yA, pA);
        if (m + kThreads >= N) break;
        fetch(m + 2 * kThreads, yA, pA);
        column(m + kThreads, yB, pB);
    }
    __shared__ double sh[kThreads / kWave];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

}  // namespace fitsne
