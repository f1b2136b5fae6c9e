// fitsne_kernels.hpp -- HIP kernels of the FIt-SNE per-iteration gradient (gfx950, wave64).
//
// Two paths share the FFT machinery:
//  * FP64 "plain": the reference's stages one for one -- charges [1,x,y,x^2+y^2] (or the df != 1 sets),
//    scattered spread with FP64 atomics, per-term circulant convolution, interpolation of every
//    potential, the Z expansion and the dC combine (tsne.cpp:1027-1169, nbodyfft.cpp:73-248).
//  * FP32 form: exact fixed-point spread of node-relative charges (A, b, c) over runs of same-box
//    points, FFT convolution with the kernel and its odd moments, Z by Parseval in FP64,
//    interpolation fused with the optimiser update.
#pragma once

#include <cfloat>

#include "fitsne_common.hpp"
#include "fitsne_fft.hpp"

namespace fitsne {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- helpers

// Consecutive block indices land on different XCDs (round-robin dispatch over 8 XCDs); map them so
// that each XCD works on a contiguous range of rows/columns and its L2 merges the partial lines.
__device__ __forceinline__ int xcd_row(int bid, int nrows) {
    const int per = (nrows + 7) / 8;
    return (bid % 8) * per + bid / 8;
}

template <typename T>
__device__ __forceinline__ T kernel_value(int kind, double r2, double df) {
    double v;
    if (kind == 0) {
        double q = 1.0 / (1.0 + r2);
        v = q * q;                            // squared_cauchy (tsne.cpp:69-71, 84-86)
    } else if (kind == 1) {
        v = df == 1.0 ? 1.0 / (1.0 + r2) : pow(1.0 + r2 / df, -df);          // general_kernel
    } else {
        v = pow(1.0 + r2 / df, -(df + 1.0));  // squared_general_kernel
    }
    return (T)v;
}

template <typename TV, int D>
__device__ __forceinline__ void load_pt(const TV *__restrict__ Y, long i, TV (&y)[D]) {
    if constexpr (D == 2) {
        using V2 = typename std::conditional<std::is_same<TV, float>::value, float2, double2>::type;
        V2 v = reinterpret_cast<const V2 *>(Y)[i];
        y[0] = v.x;
        y[1] = v.y;
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) y[d] = Y[i * D + d];
    }
}

// zeroMean applied on load: f_d(y) = (float)((double)y - mean_d), exactly what k_center stores, so a
// reader of a Y whose centring is still pending (the mean in ctr) sees the same floats; with mean 0 it is
// the identity.
__device__ __forceinline__ float centred(float y, double m) { return (float)((double)y - m); }

// Copy the twiddle table into LDS behind the workgroup's lines (the FFT stages read it once per
// butterfly input); visible after the caller's next __syncthreads.
template <typename T>
__device__ __forceinline__ const cx<T> *stage_twiddles(cx<T> *lds, const cx<T> *__restrict__ tw, int n) {
    for (int e = threadIdx.x; e < n; e += blockDim.x) lds[e] = tw[e];
    return lds;
}

// ---------------------------------------------------------------- twiddles

template <typename T>
__global__ void k_twiddles(cx<T> *tw, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) {
        double s, c;
        sincospi(-2.0 * (double)e / (double)n, &s, &c);
        tw[e] = cx<T>{(T)c, (T)s};
    }
}

// ---------------------------------------------------------------- attractive forces
// pos_i = scale * sum_j p_ij q_ij (y_i - y_j), q_ij = 1 / (1 + |y_i - y_j|^2 / df)
// (tsne.cpp:1121-1137, :965-980, :826-837, :721-732).  SW lanes per CSR row; each lane issues U
// independent (column, value) loads before its U neighbour gathers, so a wave keeps U times more
// memory requests in flight.  Column indices and values are streamed once (non-temporal loads).
template <typename TV, typename TI, typename TC, int D, int SW>
__global__ __launch_bounds__(kThreads) void k_attractive(const TI *__restrict__ rowP, const TC *__restrict__ colP,
                                                         const TV *__restrict__ valP, const TV *__restrict__ Y,
                                                         int N, TV inv_df, int df_one, TV scale,
                                                         TV *__restrict__ pos) {
    constexpr int U = 4;
    // blocks of consecutive rows go to the same XCD, so each XCD's L2 holds the Y rows its
    // neighbours gather (the points are kept in spatial order)
    constexpr int RPB = kThreads / SW;
    const long rb = (long)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const long row = rb * RPB + threadIdx.x / SW;
    const int lane = threadIdx.x % SW;
    if (row >= N) return;  // uniform across the SW lanes of a row
    const long beg = (long)rowP[row], end = (long)rowP[row + 1];
    TV yi[D];
    load_pt<TV, D>(Y, row, yi);
    TV acc[D];
#pragma unroll
    for (int d = 0; d < D; d++) acc[d] = (TV)0;
    for (long k0 = beg + lane; k0 < end; k0 += U * SW) {
        long j[U];
        TV v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long k = k0 + (long)u * SW;
            const bool ok = k < end;
            j[u] = ok ? (long)__builtin_nontemporal_load(&colP[k]) : row;  // row itself contributes 0
            v[u] = ok ? __builtin_nontemporal_load(&valP[k]) : (TV)0;
        }
        TV yj[U][D];
#pragma unroll
        for (int u = 0; u < U; u++) load_pt<TV, D>(Y, j[u], yj[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            TV diff[D], d2 = (TV)0;
#pragma unroll
            for (int d = 0; d < D; d++) {
                diff[d] = yi[d] - yj[u][d];
                d2 += diff[d] * diff[d];
            }
            const TV q = df_one ? (TV)1 / ((TV)1 + d2) : (TV)1 / ((TV)1 + d2 * inv_df);
            const TV vq = v[u] * q;
#pragma unroll
            for (int d = 0; d < D; d++) acc[d] += vq * diff[d];
        }
    }
#pragma unroll
    for (int d = 0; d < D; d++) acc[d] = group_sum<TV, SW>(acc[d]);
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < D; d++) pos[row * D + d] = acc[d] * scale;
    }
}

// ---------------------------------------------------------------- attractive forces, windowed (FP32 form)
// For points in spatial order the neighbours of a row lie in a few ranges of columns, but they are
// scattered inside them, so the CSR kernel's gathers are one L2 request per entry.  This layout works
// from LDS instead.  Rows are cut into blocks of R (<= kAttRowsD<D>) consecutive points and columns into
// windows of kWin points.  The entries of row block b whose column falls in window w form bucket (b, w);
// buckets are stored contiguously, b-major then w, each entry as
//   key = (row - b R) << 16 | (col - w kWin)   (u32)   and   p_ij (f32).
// A bucket is a pure section of row-pure quads (4 consecutive entries of one row) followed by a mixed
// section padded to a multiple of 4 entries (padding: key 0, value 0).  Each 64-entry chunk of a row's
// CSR range is split by window into groups; a group of n entries puts 4 floor(n / 4) of them in the
// pure section and the other n mod 4 in the mixed one.
// The workgroup of row block b holds its rows' Y and one int64 accumulator per row and coordinate in
// LDS, and walks its non-empty buckets: window w of Y sits in LDS (the next one is loaded into registers
// meanwhile), every lane takes 4 consecutive entries per 16-byte load, gathers both points from LDS and
// adds
//   round(p_ij q_ij (y_i - y_j) S)
// to the row's accumulator (LDS 64-bit atomics).  A pure quad reads its row once and issues one atomic
// per coordinate; a mixed quad merges its consecutive entries of one row and issues one per run.  Most
// entries are in pure quads, so the per-run atomics of a wave (issued whenever any lane changes row)
// are confined to the short mixed sections.
// S = 2^(50 - e), 2^e > max over all rows of sum_j p_ij * sqrt(df)/2 (bmax[NB], the maximum of the row
// blocks' maxima), bounds every partial sum by 2^50 since |q d| <= sqrt(df)/2: the integer sums are
// exact, so the forces depend neither on the order of the entries nor on the row-block size R, and are
// bitwise repeatable.  Same formula as k_attractive (tsne.cpp:1121-1137).
constexpr int kWin = 8192;
constexpr int kAttRows = 4096;
constexpr int kAttRows1 = 8192;  // 1D: its smaller LDS record per row allows blocks twice as long
template <int D>
constexpr int kAttRowsD = D == 1 ? kAttRows1 : kAttRows;
constexpr int kAttThreads = 1024;
// A bucket of fewer quads is processed from a flat list before the windows, with its columns' Y read
// from global memory: staging a whole window (64 KB and two barriers) for a few entries cost more.
constexpr int kSmallQuads = 512;
static_assert((kWin * 2 - 1) / 2 * 2 + 2 <= kWin * 2 && (kWin - 1) / 2 * 2 + 2 <= kWin, "small-bucket list fits");

// round-to-nearest int64 of x for |x| < 2^51: adding 1.5 * 2^52 leaves the integer in the low mantissa bits
__device__ __forceinline__ long long fix_rn(double x) {
    constexpr double M = 6755399441055744.0;
    return __double_as_longlong(x + M) - __double_as_longlong(M);
}

// fix_rn(x * s) with the scaling folded into the rounding add (one FMA); |x * s| < 2^51.
__device__ __forceinline__ long long fix_rn_fma(double x, double s) {
    constexpr double M = 6755399441055744.0;
    return __double_as_longlong(fma(x, s, M)) - __double_as_longlong(M);
}

// Round to nearest (ties to even, as fix_rn) for any |x| < 2^63: fix_rn inside its range, the full
// conversion outside it (points that the reference's min/max rule leaves outside the grid are
// extrapolated and can carry large weights).
__device__ __forceinline__ long long fix_rn_any(double x) {
    if (fabs(x) < 2251799813685248.0) return fix_rn(x);  // 2^51
    return (long long)rint(x);
}

// The window groups of one 64-entry chunk of a row (entries k0 + lane < k1), in a fixed order: the
// wave-uniform loop takes the window of the lowest remaining lane, the lanes that share it form the
// group.  f(w, m) runs on every lane for each group (m: its lanes).
template <typename F>
__device__ __forceinline__ void window_groups(bool valid, int w, F &&f) {
    unsigned long long left = __ballot(valid);
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const int wl = __shfl(w, leader, kWave);
        const unsigned long long m = __ballot(valid && w == wl);
        f(wl, m, leader);
        left &= ~m;
    }
}

// Per row block: the size of every bucket (pure entries + mixed entries rounded up to a multiple of 4)
// into count[b * NW + w], its pure entries into pure[b * NW + w], and max_i sum_j |p_ij| into bmax[b].
__global__ __launch_bounds__(kAttThreads) void k_win_count(const int64_t *__restrict__ rowP,
                                                           const uint32_t *__restrict__ colP,
                                                           const float *__restrict__ val, int N, int R, int NW,
                                                           const int *__restrict__ rstart,
                                                           int *__restrict__ count, int *__restrict__ pure,
                                                           float *__restrict__ bmax) {
    extern __shared__ int hist[];  // [0, NW): pure entries, [NW, 2 NW): mixed entries
    __shared__ unsigned int smax;
    const int b = blockIdx.x, t = threadIdx.x, lane = t % kWave, wv = t / kWave;
    // row block b: [b R, (b + 1) R), or [rstart[b], rstart[b + 1]) for a balanced partition
    const int r0 = rstart ? rstart[b] : b * R, r1 = rstart ? rstart[b + 1] : min(N, r0 + R);
    for (int w = t; w < 2 * NW; w += kAttThreads) hist[w] = 0;
    if (t == 0) smax = 0u;
    __syncthreads();
    for (int r = r0 + wv; r < r1; r += kAttThreads / kWave) {
        float sum = 0.f;
        const long k1 = rowP[r + 1];
        for (long k0 = rowP[r]; k0 < k1; k0 += kWave) {
            const long k = k0 + lane;
            const bool valid = k < k1;
            const int w = valid ? (int)(colP[k] / kWin) : -1;
            if (valid) sum += fabsf(val[k]);
            window_groups(valid, w, [&](int wl, unsigned long long m, int leader) {
                if (lane == leader) {
                    const int n = __popcll(m);
                    atomicAdd(&hist[wl], n & ~3);
                    atomicAdd(&hist[NW + wl], n & 3);
                }
            });
        }
        sum = group_sum<float, kWave>(sum);
        if (lane == 0) atomicMax(&smax, __float_as_uint(sum));  // non-negative floats order as integers
    }
    __syncthreads();
    for (int w = t; w < NW; w += kAttThreads) {
        count[(long)b * NW + w] = hist[w] + ((hist[NW + w] + 3) & ~3);
        pure[(long)b * NW + w] = hist[w];
    }
    if (t == 0) bmax[b] = __uint_as_float(smax);
}

// bmax[nb] = max of bmax[0 .. nb) (one workgroup of 1024; non-negative floats order as their bits).
__global__ __launch_bounds__(1024) void k_bmax_all(float *__restrict__ bmax, int nb) {
    __shared__ unsigned int m;
    if (threadIdx.x == 0) m = 0u;
    __syncthreads();
    unsigned int v = 0u;
    for (int b = threadIdx.x; b < nb; b += 1024) v = max(v, __float_as_uint(bmax[b]));
    atomicMax(&m, v);
    __syncthreads();
    if (threadIdx.x == 0) bmax[nb] = __uint_as_float(m);
}

// Scatter the entries into their buckets (start[] from the exclusive scan of the counts) and write the
// padding.  The groups are those of k_win_count; a group reserves its pure and mixed slots with one
// LDS atomic each, so its pure quads are row-pure.  The order of the groups inside a bucket is whatever
// the cursors give; the sums do not depend on it.
__global__ __launch_bounds__(kAttThreads) void k_win_fill(const int64_t *__restrict__ rowP,
                                                          const uint32_t *__restrict__ colP,
                                                          const float *__restrict__ val, int N, int R, int NW,
                                                          const int *__restrict__ rstart,
                                                          const int *__restrict__ start,
                                                          const int *__restrict__ pure, uint32_t *__restrict__ key,
                                                          float *__restrict__ wval) {
    extern __shared__ int cur[];  // [0, NW): pure cursors, [NW, 2 NW): mixed cursors
    const int b = blockIdx.x, t = threadIdx.x, lane = t % kWave, wv = t / kWave;
    const int r0 = rstart ? rstart[b] : b * R, r1 = rstart ? rstart[b + 1] : min(N, r0 + R);
    for (int w = t; w < 2 * NW; w += kAttThreads) cur[w] = 0;
    __syncthreads();
    const int *st = start + (long)b * NW;
    const int *pu = pure + (long)b * NW;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = r0 + wv; r < r1; r += kAttThreads / kWave) {
        const long k1 = rowP[r + 1];
        for (long k0 = rowP[r]; k0 < k1; k0 += kWave) {
            const long k = k0 + lane;
            const bool valid = k < k1;
            const uint32_t j = valid ? colP[k] : 0u;
            const int w = valid ? (int)(j / kWin) : -1;
            window_groups(valid, w, [&](int wl, unsigned long long m, int leader) {
                const int n = __popcll(m), np = n & ~3;
                int bp = 0, bm = 0;
                if (lane == leader) {
                    bp = atomicAdd(&cur[wl], np);
                    bm = atomicAdd(&cur[NW + wl], n & 3);
                }
                bp = __shfl(bp, leader, kWave);
                bm = __shfl(bm, leader, kWave);
                if (valid && w == wl) {
                    const int rank = __popcll(m & below);
                    const int at = rank < np ? st[wl] + bp + rank : st[wl] + pu[wl] + bm + (rank - np);
                    key[at] = (uint32_t)(r - r0) << 16 | (j - (uint32_t)wl * kWin);
                    wval[at] = val[k];
                }
            });
        }
    }
    __syncthreads();
    for (int w = t; w < NW; w += kAttThreads)
        for (int at = st[w] + pu[w] + cur[NW + w]; at < st[w + 1]; at++) {
            key[at] = 0u;
            wval[at] = 0.f;
        }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Four entries of one row (a quad of the pure section).
// SH: yw is the uncentred Y in global memory, centred on load by (m0, m1).
template <int D, bool SH = false>
__device__ __forceinline__ void attr_quad_pure(const u32x4 &kq, const f32x4 &vq4, const float *yr, const float *yw,
                                               unsigned long long *acc, float inv_df, int df_one, double S,
                                               double m0 = 0.0, double m1 = 0.0) {
    const uint32_t kk[4] = {kq.x, kq.y, kq.z, kq.w};
    const float v[4] = {vq4.x, vq4.y, vq4.z, vq4.w};
    const int il = (int)(kk[0] >> 16);
    float yi[D];
    long long run[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        yi[d] = yr[il * D + d];
        run[d] = 0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int jl = (int)(kk[u] & 0xffffu);
        float diff[D], d2 = 0.f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const float yj = SH ? centred(yw[jl * D + d], d ? m1 : m0) : yw[jl * D + d];
            diff[d] = yi[d] - yj;
            d2 += diff[d] * diff[d];
        }
        const float q = __builtin_amdgcn_rcpf(df_one ? 1.f + d2 : 1.f + d2 * inv_df);
        const float vq = v[u] * q;
#pragma unroll
        for (int d = 0; d < D; d++) run[d] += fix_rn_fma((double)(vq * diff[d]), S);
    }
#pragma unroll
    for (int d = 0; d < D; d++) atomicAdd(&acc[il * D + d], (unsigned long long)run[d]);
}

// Four consecutive entries of the mixed section (any rows).
template <int D, bool SH = false>
__device__ __forceinline__ void attr_quad(const u32x4 &kq, const f32x4 &vq4, const float *yr, const float *yw,
                                          unsigned long long *acc, float inv_df, int df_one, double S,
                                          double m0 = 0.0, double m1 = 0.0) {
    const uint32_t kk[4] = {kq.x, kq.y, kq.z, kq.w};
    const float v[4] = {vq4.x, vq4.y, vq4.z, vq4.w};
    int cur = -1;
    long long run[D];
#pragma unroll
    for (int d = 0; d < D; d++) run[d] = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int il = (int)(kk[u] >> 16), jl = (int)(kk[u] & 0xffffu);
        float diff[D], d2 = 0.f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const float yj = SH ? centred(yw[jl * D + d], d ? m1 : m0) : yw[jl * D + d];
            diff[d] = yr[il * D + d] - yj;
            d2 += diff[d] * diff[d];
        }
        // hardware reciprocal (1 ulp): well inside the FP32 form's error, a tenth of a division's cost
        const float q = __builtin_amdgcn_rcpf(df_one ? 1.f + d2 : 1.f + d2 * inv_df);
        const float vq = v[u] * q;
        if (il != cur) {
            if (cur >= 0) {
#pragma unroll
                for (int d = 0; d < D; d++) atomicAdd(&acc[cur * D + d], (unsigned long long)run[d]);
            }
            cur = il;
#pragma unroll
            for (int d = 0; d < D; d++) run[d] = 0;
        }
#pragma unroll
        for (int d = 0; d < D; d++) run[d] += fix_rn_fma((double)(vq * diff[d]), S);
    }
#pragma unroll
    for (int d = 0; d < D; d++) atomicAdd(&acc[cur * D + d], (unsigned long long)run[d]);
}

// LDS of one k_attr_win workgroup: one workgroup per CU, 4 waves per SIMD (160 KB in 2D; 128 KB in 1D,
// whose blocks are twice as long so that N = 2M still needs a single round of workgroups).
template <int D>
struct AttrLds {
    float yw[kWin * D];                   // the staged window's Y
    float yr[kAttRowsD<D> * D];               // the row block's Y
    unsigned long long acc[kAttRowsD<D> * D];  // the row block's fixed-point sums
};

// Row block b of the windowed attractive term (the body of k_attr_win, and of its batched form).
template <int D>
__device__ __forceinline__ void attr_win_block(const int *__restrict__ start, const int *__restrict__ pure,
                                               const uint32_t *__restrict__ key, const float *__restrict__ wval,
                                               const float *__restrict__ bmax, const float *__restrict__ Y, int N,
                                               int R, int NW, const int *__restrict__ rstart, int NB, float inv_df,
                                               int df_one, float mdf, float scale, float *__restrict__ pos, int b,
                                               AttrLds<D> &L, const double *__restrict__ ctr = nullptr) {
    constexpr int WL = kWin * D / kAttThreads;  // window floats per thread
    float *yw = L.yw, *yr = L.yr;
    unsigned long long *acc = L.acc;
    const int t = threadIdx.x;
    // a pending centring (ctr) is applied as Y is staged; a thread's flat indices q = t + u kAttThreads
    // all hold coordinate t mod D
    const bool sh = ctr != nullptr;
    const double m0 = sh ? ctr[0] : 0.0, m1 = D == 2 && sh ? ctr[1] : 0.0;
    const int r0 = rstart ? rstart[b] : b * R, nr = rstart ? rstart[b + 1] - r0 : min(R, N - r0);
    {
        const double mt = D == 2 && (t & 1) ? m1 : m0;
        for (int q = t; q < nr * D; q += kAttThreads) {
            yr[q] = sh ? centred(Y[(long)r0 * D + q], mt) : Y[(long)r0 * D + q];
            acc[q] = 0ull;
        }
    }
    const double bound = (double)bmax[NB] * mdf;  // the maximum over all NB row blocks
    int e = 0;
    if (bound > 0.0) (void)frexp(bound, &e);  // bound < 2^e
    const double S = ldexp(1.0, 50 - e), invS = ldexp(1.0, e - 50);
    const int *st = start + (long)b * NW;
    const int *pu = pure + (long)b * NW;
    const u32x4 *K4 = reinterpret_cast<const u32x4 *>(key);
    const f32x4 *V4 = reinterpret_cast<const f32x4 *>(wval);
    // Small buckets: listed (window numbers, then the exclusive prefix of their quads) in the window
    // buffer, which is free until the first window is staged, and processed as one flat range of quads.
    // If more than fit are found, none is listed and every bucket is staged.
    constexpr int kSmCap = (kWin * D - 1) / 2;  // sm_w[kSmCap], sm_pre[kSmCap + 1], count: kWin * D ints
    int *sm_w = reinterpret_cast<int *>(yw);
    int *sm_pre = sm_w + kSmCap;
    int *sm_n = sm_pre + kSmCap + 1;
    if (t == 0) *sm_n = 0;
    __syncthreads();  // also: yr and acc are initialised
    for (int ww = t; ww < NW; ww += kAttThreads) {
        const int nq = (st[ww + 1] - st[ww]) >> 2;
        if (nq > 0 && nq < kSmallQuads) {
            const int slot = atomicAdd(sm_n, 1);
            if (slot < kSmCap) sm_w[slot] = ww;
        }
    }
    __syncthreads();
    const int ns = *sm_n;
    const bool flat = ns <= kSmCap;
    if (flat && ns > 0) {
        if (t < kWave) {  // one wave: exclusive prefix of the listed buckets' quads (list order)
            int carry = 0;
            for (int i0 = 0; i0 < ns; i0 += kWave) {
                const int i = i0 + t;
                const int nq = i < ns ? (st[sm_w[i] + 1] - st[sm_w[i]]) >> 2 : 0;
                int incl = nq;
                for (int o = 1; o < kWave; o <<= 1) {
                    const int y = __shfl_up(incl, o, kWave);
                    if (t >= o) incl += y;
                }
                if (i < ns) sm_pre[i] = carry + incl - nq;
                carry += __shfl(incl, kWave - 1, kWave);
            }
            if (t == 0) sm_pre[ns] = carry;
        }
        __syncthreads();
        const int total = sm_pre[ns];
        for (int qi = t; qi < total; qi += kAttThreads) {
            int lo = 0, hi = ns - 1;  // the listed bucket holding quad qi (prefixes strictly increase)
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sm_pre[mid] <= qi) lo = mid;
                else hi = mid - 1;
            }
            const int ww = sm_w[lo];
            const int qq = (st[ww] >> 2) + (qi - sm_pre[lo]);
            const int qm = (st[ww] + pu[ww]) >> 2;
            const float *ywg = Y + (long)ww * kWin * D;  // the window's Y in global memory
            const u32x4 ka = __builtin_nontemporal_load(&K4[qq]);
            const f32x4 va = __builtin_nontemporal_load(&V4[qq]);
            if (sh) {  // uniform: the centring's FP64 work only when one is pending
                if (qq < qm) attr_quad_pure<D, true>(ka, va, yr, ywg, acc, inv_df, df_one, S, m0, m1);
                else attr_quad<D, true>(ka, va, yr, ywg, acc, inv_df, df_one, S, m0, m1);
            } else {
                if (qq < qm) attr_quad_pure<D>(ka, va, yr, ywg, acc, inv_df, df_one, S);
                else attr_quad<D>(ka, va, yr, ywg, acc, inv_df, df_one, S);
            }
        }
    }
    __syncthreads();  // the list lives in the window buffer
    auto staged = [&](int ww) {
        const int nq = (st[ww + 1] - st[ww]) >> 2;
        return nq > 0 && !(flat && nq < kSmallQuads);
    };
    int w = 0;
    while (w < NW && !staged(w)) w++;
    // the window's loads stay in flight across the current bucket; a pending centring is applied as
    // they are stored to LDS
    float pre[WL];
    auto fetch = [&](int ww) {
        const long base = (long)ww * kWin * D;
        const int lim = (int)min((long)kWin * D, (long)N * D - base);
#pragma unroll
        for (int u = 0; u < WL; u++) {
            const int q = t + u * kAttThreads;
            pre[u] = q < lim ? Y[base + q] : 0.f;
        }
    };
    // Each thread takes quads qq and qq + kAttThreads of a round.  It loads the next round's pair (in
    // the same bucket, or the first pair of the next staged bucket) before it computes, so that two
    // pairs per thread are in flight across the compute and across the barriers of a window change.
    u32x4 ka = {0u, 0u, 0u, 0u}, kb = {0u, 0u, 0u, 0u};
    f32x4 va = {0.f, 0.f, 0.f, 0.f}, vb = {0.f, 0.f, 0.f, 0.f};
    auto load_pair = [&](int qq, int q1) {
        if (qq < q1) {
            ka = __builtin_nontemporal_load(&K4[qq]);
            va = __builtin_nontemporal_load(&V4[qq]);
        }
        if (qq + kAttThreads < q1) {
            kb = __builtin_nontemporal_load(&K4[qq + kAttThreads]);
            vb = __builtin_nontemporal_load(&V4[qq + kAttThreads]);
        }
    };
    if (w < NW) {
        fetch(w);
        load_pair((st[w] >> 2) + t, st[w + 1] >> 2);
    }
    while (w < NW) {
        const int q0 = st[w] >> 2, qm = (st[w] + pu[w]) >> 2, q1 = st[w + 1] >> 2;
        int wn = w + 1;
        while (wn < NW && !staged(wn)) wn++;
        __syncthreads();  // the previous window's readers are done
        if (sh) {
            // the lane's mean, read here (after the barrier) so that it is not held across the loop;
            // the padding past N is centred too and never read
            const double mt = ctr[D == 2 ? (t & 1) : 0];
#pragma unroll
            for (int u = 0; u < WL; u++) pre[u] = centred(pre[u], mt);
        }
#pragma unroll
        for (int u = 0; u < WL; u++) yw[t + u * kAttThreads] = pre[u];
        __syncthreads();
        if (wn < NW) fetch(wn);  // consumed after this bucket
        int qq = q0 + t;
        if (qq >= q1 && wn < NW) load_pair((st[wn] >> 2) + t, st[wn + 1] >> 2);  // no quad here
        for (; qq < q1; qq += 2 * kAttThreads) {
            const u32x4 ca = ka, cb = kb;
            const f32x4 cva = va, cvb = vb;
            const int qn = qq + 2 * kAttThreads;
            if (qn < q1) load_pair(qn, q1);
            else if (wn < NW) load_pair((st[wn] >> 2) + t, st[wn + 1] >> 2);
            if (qq < qm) attr_quad_pure<D>(ca, cva, yr, yw, acc, inv_df, df_one, S);
            else attr_quad<D>(ca, cva, yr, yw, acc, inv_df, df_one, S);
            if (qq + kAttThreads < q1) {
                if (qq + kAttThreads < qm) attr_quad_pure<D>(cb, cvb, yr, yw, acc, inv_df, df_one, S);
                else attr_quad<D>(cb, cvb, yr, yw, acc, inv_df, df_one, S);
            }
        }
        w = wn;
    }
    __syncthreads();
    for (int q = t; q < nr * D; q += kAttThreads)
        pos[(long)r0 * D + q] = (float)((double)(long long)acc[q] * invS) * scale;
}

template <int D>
__global__ __launch_bounds__(kAttThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_attr_win(const int *__restrict__ start,
                                                          const int *__restrict__ pure,
                                                          const uint32_t *__restrict__ key,
                                                          const float *__restrict__ wval,
                                                          const float *__restrict__ bmax,
                                                          const float *__restrict__ Y, int N, int R, int NW,
                                                          const int *__restrict__ rstart, int nblk,
                                                          float inv_df, int df_one, float mdf,
                                                          const StepDev *__restrict__ sd, float exag,
                                                          const DevGeo *__restrict__ dg, float *__restrict__ pos,
                                                          const double *__restrict__ ctr) {
    if (!dg->live) return;
    const float scale = exag >= 0.f ? exag : sd->exag;  // exag < 0: the step record's (captured steps)
    __shared__ AttrLds<D> L;
    const int b = xcd_row(blockIdx.x, nblk);
    if (b >= nblk) return;
    attr_win_block<D>(start, pure, key, wval, bmax, Y, N, R, NW, rstart, nblk, inv_df, df_one, mdf, scale, pos, b,
                      L, ctr);
}

// ---------------------------------------------------------------- min / max
// Over the flat sequence Y[0..n): max of all; min of all but index 0 (2D) or of all (1D); Y[0].
// The 2D reference rule (`else if`, tsne.cpp:1045-1048) equals this whenever Y[0] >= min(rest);
// otherwise the host runs k_minmax_seq.
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_minmax_partial(const TV *__restrict__ Y, long n, long skip,
                                                             double *__restrict__ part) {
    double mx = -INFINITY, mn = INFINITY;
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads) {
        const double v = (double)Y[e];
        mx = fmax(mx, v);
        if (e != skip) mn = fmin(mn, v);
    }
    __shared__ double sh[2][kThreads / kWave];
    mx = wave_max(mx);
    mn = wave_min(mn);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        sh[0][w] = mx;
        sh[1][w] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kThreads / kWave; i++) {
            mx = fmax(mx, sh[0][i]);
            mn = fmin(mn, sh[1][i]);
        }
        part[blockIdx.x * 2 + 0] = mx;
        part[blockIdx.x * 2 + 1] = mn;
    }
}

template <typename TV>
__global__ void k_minmax_final(const double *__restrict__ part, int nblk, const TV *__restrict__ Y, long first,
                               double *__restrict__ out) {
    double mx = -INFINITY, mn = INFINITY;
    const int T = blockDim.x;
    int b = threadIdx.x;
    for (; b + 7 * T < nblk; b += 8 * T) {  // loads issued eight at a time
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = reinterpret_cast<const double2 *>(part)[b + u * T];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            mx = fmax(mx, v[u].x);
            mn = fmin(mn, v[u].y);
        }
    }
    for (; b < nblk; b += T) {
        mx = fmax(mx, part[2 * b]);
        mn = fmin(mn, part[2 * b + 1]);
    }
    mx = wave_max(mx);
    mn = wave_min(mn);
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = mn;
        out[2] = (double)Y[first];
    }
}

// Exact sequential `if (v > max) max = v; else if (v < min) min = v;` over Y[0..n) in one block:
// chunk maxima, their exclusive prefix maximum, then each chunk replays the rule from its prefix.
// ctr: a pending centring of the FP32 Y (2D), applied on load
template <typename TV>
__device__ __forceinline__ double seq_at(const TV *__restrict__ Y, const int *__restrict__ pos, long e,
                                         const double *__restrict__ ctr) {
    const TV v = pos ? Y[2L * pos[e >> 1] + (e & 1)] : Y[e];
    if constexpr (std::is_same<TV, float>::value) {
        if (ctr) return (double)centred(v, ctr[e & 1]);
    }
    return (double)v;
}

// One block of 1024 threads; out[0] = max, out[1] = min.
template <typename TV>
__device__ void minmax_seq_block(const TV *__restrict__ Y, const int *__restrict__ pos, long n, double *out,
                                 const double *__restrict__ ctr = nullptr) {
    __shared__ double cmax[1024];
    __shared__ double cmin[1024];
    const int t = threadIdx.x;
    const long chunk = (n + 1023) / 1024;
    const long b = t * chunk, e = b + chunk < n ? b + chunk : n;
    double m = -INFINITY;
    for (long i = b; i < e; i++) m = fmax(m, seq_at(Y, pos, i, ctr));
    cmax[t] = m;
    __syncthreads();
    if (t == 0) {
        double run = -INFINITY;
        for (int i = 0; i < 1024; i++) {
            double c = cmax[i];
            cmax[i] = run;  // exclusive prefix
            run = fmax(run, c);
        }
        out[0] = run;
    }
    __syncthreads();
    double mx = cmax[t], mn = INFINITY;
    for (long i = b; i < e; i++) {
        const double v = seq_at(Y, pos, i, ctr);
        if (v > mx) mx = v;
        else if (v < mn) mn = v;
    }
    cmin[t] = mn;
    __syncthreads();
    if (t == 0) {
        double r = INFINITY;
        for (int i = 0; i < 1024; i++) r = fmin(r, cmin[i]);
        out[1] = r;
    }
    __syncthreads();
}

template <typename TV>
__global__ __launch_bounds__(1024) void k_minmax_seq(const TV *__restrict__ Y, const int *__restrict__ pos, long n,
                                                     double *__restrict__ out, const double *__restrict__ ctr) {
    minmax_seq_block(Y, pos, n, out, ctr);
}

// ---------------------------------------------------------------- step setup on the device
// The geometry of a captured step, from the min / max partials the previous step (or the upload of Y)
// left: max and min(rest) from the partials; when Y[first] is the strict minimum the sequential
// 2D rule is replayed (tsne.cpp:1045-1048); then the box count, geometry and plan of make_geo /
// make_plan, the refusal of grids beyond n1d_max / mp_max (live = 0, err = 1) or of a step whose
// spread has no usable fixed-point scale (err = 2, spread_fix_scale), and whether the kernel spectra
// and twiddles of the previous step still apply.  One block of 1024 threads.
// The body of k_geometry (and of its batched form); one workgroup of 1024.  When tw is given, the same
// workgroup then writes the twiddle table of the step's FFT length if it changed (k_twiddles_dev's values).
template <typename TV>
__device__ void geometry_block(const double *__restrict__ part, int nblk, const TV *__restrict__ Y, long first,
                               const int *__restrict__ pos, long n, int D, int p, double ipi, int mni, int n1d_max,
                               int mp_max, double lb, DevGeo *__restrict__ dg, cx<float> *__restrict__ tw) {
    __shared__ double shx[1024 / kWave], shn[1024 / kWave];
    __shared__ double res[2];
    __shared__ int twn;
    double mx = -INFINITY, mn = INFINITY;
    for (int b = threadIdx.x; b < nblk; b += 1024) {
        mx = fmax(mx, part[2 * b]);
        mn = fmin(mn, part[2 * b + 1]);
    }
    mx = wave_max(mx);
    mn = wave_min(mn);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        shx[w] = mx;
        shn[w] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 1024 / kWave; i++) {
            mx = fmax(mx, shx[i]);
            mn = fmin(mn, shn[i]);
        }
        res[0] = mx;
        res[1] = mn;
    }
    __syncthreads();
    mx = res[0];
    mn = res[1];
    if (D == 2 && (double)Y[first] < mn) {  // uniform over the block
        __syncthreads();
        minmax_seq_block(Y, pos, n, res);
        mx = res[0];
        mn = res[1];
    }
    if (threadIdx.x == 0) {
        DevGeo o = *dg;
        o.mn = mn;
        o.mx = mx;
        const bool fin = isfinite(mn) && isfinite(mx);
        const int nb = fin ? (D == 2 ? n_boxes_2d(mn, mx, ipi, mni) : n_boxes_1d(mn, mx, ipi, mni)) : 0;
        Geo g{};
        if (nb >= 1) g = make_geo(D, p, mn, mx, nb);
        o.live = fin && nb >= 1 && g.n1d <= n1d_max && g.Mp <= mp_max;
        if (o.live) {  // the spread's fixed-point scale, with the first coordinate's gap below the grid
            const double y0 = (double)Y[first];
            const double gap = D == 2 && y0 < mn ? (mn - y0) / g.bw : 0.0;
            g.fix = spread_fix_scale(g, (double)(n / D), gap, lb);
            g.fix_fast = spread_fix_fast(g, gap, lb);
            if (g.fix == 0) {
                o.live = 0;
                o.err = 2;
            }
        } else {
            o.err = 1;
        }
        if (o.live) {
            o.g = g;
            o.pl = make_plan(g.Mp);
            o.kspec_stale = !(o.kc_nb == nb && o.kc_mn == mn && o.kc_mx == mx);
            o.kc_nb = nb;
            o.kc_mn = mn;
            o.kc_mx = mx;
            o.tw_stale = o.tw_n != g.Mp;
            o.tw_n = g.Mp;
        }
        *dg = o;
        twn = o.live && o.tw_stale ? g.Mp : 0;
    }
    if (tw) {
        __syncthreads();
        const int tn = twn;
        for (int e = threadIdx.x; e < tn; e += blockDim.x) {
            double sn, cs;
            sincospi(-2.0 * (double)e / (double)tn, &sn, &cs);
            tw[e] = cx<float>{(float)cs, (float)sn};
        }
    }
}

template <typename TV>
__global__ __launch_bounds__(1024) void k_geometry(const double *__restrict__ part, int nblk, const TV *__restrict__ Y,
                                                   long first, const int *__restrict__ pos, long n, int D, int p,
                                                   double ipi, int mni, int n1d_max, int mp_max, double lb,
                                                   DevGeo *__restrict__ dg) {
    geometry_block<TV>(part, nblk, Y, first, pos, n, D, p, ipi, mni, n1d_max, mp_max, lb, dg, nullptr);
}

// Twiddle table of the step's FFT length, when it changed.
template <typename T>
__global__ void k_twiddles_dev(const DevGeo *__restrict__ dg, cx<T> *__restrict__ tw) {
    if (!dg->live || !dg->tw_stale) return;
    const int n = dg->g.Mp;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        double sn, cs;
        sincospi(-2.0 * (double)e / (double)n, &sn, &cs);
        tw[e] = cx<T>{(T)cs, (T)sn};
    }
}

// Host-computed geometry / step parameters into device memory (stream-ordered, no host buffer kept).
__global__ void k_set_geo(DevGeo v, DevGeo *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = v;
}
__global__ void k_set_step(StepDev v, StepDev *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = v;
}
__global__ void k_clear_err(DevGeo *__restrict__ dg) {
    if (threadIdx.x == 0 && blockIdx.x == 0) dg->err = 0;
}

// ---------------------------------------------------------------- exclusive scan

// Exclusive scan of count[0..n) into start[0..n], then count[] = 0 for the next use.  One block of
// 1024 threads; per tile of 4096 counters each thread owns 4 consecutive ones, and the tile is scanned
// with wave shuffles plus one pass over the 16 wave totals; a carry runs across tiles.
__global__ __launch_bounds__(1024) void k_scan_counts(int *__restrict__ count, int *__restrict__ start, int n) {
    __shared__ int wsum[1024 / kWave];
    __shared__ int tile_total;
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    int carry = 0;
    for (int base = 0; base < n; base += 4 * 1024) {
        const int i0 = base + 4 * t;
        int v[4], loc = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = i0 + q < n ? count[i0 + q] : 0;
            loc += v[q];
        }
        int x = loc;  // inclusive wave scan of the per-thread totals
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(x, o, kWave);
            if (lane >= o) x += y;
        }
        if (lane == kWave - 1) wsum[w] = x;
        __syncthreads();
        if (t == 0) {
            int run = 0;
            for (int i = 0; i < 1024 / kWave; i++) {
                const int c = wsum[i];
                wsum[i] = run;
                run += c;
            }
            tile_total = run;
        }
        __syncthreads();
        int e = carry + wsum[w] + x - loc;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (i0 + q < n) {
                start[i0 + q] = e;
                count[i0 + q] = 0;
            }
            e += v[q];
        }
        carry += tile_total;
        __syncthreads();
    }
    if (t == 0) start[n] = carry;
}

// ---------------------------------------------------------------- spread, FP32 form
// A workgroup takes `pts` consecutive points (a multiple of 2 kSpreadThreads; memory order; spatially
// ordered points keep this to a few dozen boxes).  Each lane takes two consecutive points, finds their
// boxes and computes the node-relative charges
//   A = L,  c = L |e|^2,  b = L e    (e = y - node, in box widths)
// for the p^D nodes of each point's box.  Runs of consecutive points in one box are summed in FP32 (a
// lane's pair, then a segmented scan over the lanes in the VALU: DPP, a fixed order over fixed points),
// each run's sums are scaled by g.fix (2^38 unless the step's bound needs less), rounded
// to int64 and added into the workgroup's LDS table slot of its box (open addressing, LDS atomics), and
// the table goes to the grid with one 64-bit global atomic per (box, node, term).  A run whose box finds
// no free slot adds to the grid directly.  The sums across runs are exact integers, so the grid, and so
// the whole gradient, is bitwise repeatable whatever the order in which workgroups and waves run, and
// whatever `pts`: the waves are the same aligned groups of 128 points for every choice.
// Grid layout: int64 [u][v][4] (A, c, b_x, b_y), u the x node (nbodyfft.cpp:137-138); the row FFT
// reads it, converts it and clears it for the next iteration.
constexpr int kSpreadThreads = 256;
constexpr int kSpreadPts = 1024;  // default points per workgroup (the autotuner may choose another)

// One step of a segmented inclusive scan over the 64 lanes: each lane adds the value that the DPP
// control CTRL (row_shr:n, row_bcast:15 / :31 with row mask RMASK) brings from a lower lane, when that
// lane is in the same run.  Lanes without a source get run id 0, which is never a real run.
template <int CTRL, int RMASK, int NV, int NQ>
__device__ __forceinline__ void seg_scan_step(float (&val)[NV][NQ], int run) {
    const int src_run = __builtin_amdgcn_update_dpp(0, run, CTRL, RMASK, 0xf, false);
    const bool take = src_run == run;
    float y[NV][NQ];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int q = 0; q < NQ; q++)
            y[v][q] = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(val[v][q]), CTRL, RMASK, 0xf, false));
    if (take) {  // one exec mask for all the adds of the step
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int q = 0; q < NQ; q++) val[v][q] += y[v][q];
    }
}

// Table slots of the spread: the table (sums and keys) stays within 40 KB, so that four workgroups fit
// in a CU's LDS.
template <int P, int D>
constexpr int spread_slots() {
    constexpr int per = (D == 2 ? P * P * 4 : P * 3) * 8 + 4;  // bytes of one table slot: sums and key
    constexpr int s = (40 * 1024) / per;
    return s >= 256 ? 256 : s;
}

// In 2D with p <= 3 the registers are held to 128 per lane (4 waves per SIMD), so that, with the table
// under 40 KB, four workgroups share a CU (three at the 150 registers the compiler chose on its own:
// measured 0.355 against 0.362 ms per step at N = 1M, results unchanged).  For p = 4, 5 the limit would
// spill most of the kernel's values, so it is not requested.
// LDS table of one k_spread_f32 workgroup.
template <int P, int D>
struct SpreadLds {
    static constexpr int NV = (D == 2 ? P * P : P);
    static constexpr int NQ = (D == 2 ? 4 : 3);
    static constexpr int NS = spread_slots<P, D>();
    int tkey[NS];
    unsigned long long tval[NS][NV * NQ];
};

// Charges of one point at its box: the NV x NQ values of its Lagrange weights (2D: A, c, b_x, b_y per
// node; 1D: A, c, b per node), and its box number.
template <int P, int D>
__device__ __forceinline__ int point_charges(const float (&y)[D], const Geo &g, float (&val)[D == 2 ? P * P : P][D == 2 ? 4 : 3]) {
    double tx;
    int box = box_coord<D>((double)y[0], g, tx);
    float Lx[P], dx[P];
#pragma unroll
    for (int a = 0; a < P; a++) dx[a] = (float)(tx - g.sp[a]);
#pragma unroll
    for (int a = 0; a < P; a++) {
        float v = 1.f;
#pragma unroll
        for (int c = 0; c < P; c++)
            if (c != a) v *= dx[c];
        Lx[a] = v * g.rdenf[a];
    }
    if constexpr (D == 2) {
        double ty;
        box += box_coord<D>((double)y[1], g, ty) * g.nb;
        float Ly[P], dy[P];
#pragma unroll
        for (int a = 0; a < P; a++) dy[a] = (float)(ty - g.sp[a]);
#pragma unroll
        for (int a = 0; a < P; a++) {
            float v = 1.f;
#pragma unroll
            for (int c = 0; c < P; c++)
                if (c != a) v *= dy[c];
            Ly[a] = v * g.rdenf[a];
        }
#pragma unroll
        for (int a = 0; a < P; a++)
#pragma unroll
            for (int c = 0; c < P; c++) {
                const float w = Lx[a] * Ly[c];
                val[a * P + c][0] = w;
                val[a * P + c][1] = w * (dx[a] * dx[a] + dy[c] * dy[c]);
                val[a * P + c][2] = w * dx[a];
                val[a * P + c][3] = w * dy[c];
            }
    } else {
#pragma unroll
        for (int a = 0; a < P; a++) {
            val[a][0] = Lx[a];
            val[a][1] = Lx[a] * dx[a] * dx[a];
            val[a][2] = Lx[a] * dx[a];
        }
    }
    return box;
}

// Workgroup `bid` of the spread (points [bid pts, (bid + 1) pts)); the body of k_spread_f32 and of its
// batched form.  Each lane takes two consecutive points, a and b, so a wavefront covers 128 points and
// its scan and flush serve twice as many points.  A lane's running sum is that of b's box: a joins it
// when both share a box; otherwise a joins the left neighbour's sum when it continues that lane's b
// box, and is flushed on its own in the rare remaining case (a box holding a alone between two others).
// Runs of lanes with the same b box are summed by the segmented scan, and each run's last lane flushes.
template <int P, int D>
__device__ __forceinline__ void spread_block(const float *__restrict__ Y, int N, int pts, const DevGeo *__restrict__ dg,
                                             unsigned long long *__restrict__ grid, int bid, SpreadLds<P, D> &L,
                                             const double *__restrict__ ctr = nullptr) {
    constexpr int NV = SpreadLds<P, D>::NV;
    constexpr int NQ = SpreadLds<P, D>::NQ;
    constexpr int NS = SpreadLds<P, D>::NS;
    int *tkey = L.tkey;
    auto &tval = L.tval;
    const Geo &g = dg->g;
    const int t = threadIdx.x, lane = t % kWave;
    for (int o = t; o < NS; o += kSpreadThreads) tkey[o] = -1;
    for (int o = t; o < NS * NV * NQ; o += kSpreadThreads) (&tval[0][0])[o] = 0ull;
    __syncthreads();
    // flush `val` of the lanes with `on` into box `box`: its LDS table slot (open addressing), or the
    // grid when the table has no free slot (rare: a separate pass)
    auto flush = [&](bool on, int box, const float (&val)[NV][NQ]) {
        int slot = -1;
        if (on) {
            int h = box % NS;
            for (int probe = 0; probe < NS; probe++) {
                const int k = atomicCAS(&tkey[h], -1, box);
                if (k == -1 || k == box) {
                    slot = h;
                    break;
                }
                h = h + 1 == NS ? 0 : h + 1;
            }
        }
        if (on && slot >= 0) {
            if (g.fix_fast) {  // uniform: every run sum within the 2^51 range of fix_rn
#pragma unroll
                for (int v = 0; v < NV; v++)
#pragma unroll
                    for (int q = 0; q < NQ; q++)
                        atomicAdd(&tval[slot][v * NQ + q], (unsigned long long)fix_rn((double)val[v][q] * g.fix));
            } else {
#pragma unroll
                for (int v = 0; v < NV; v++)
#pragma unroll
                    for (int q = 0; q < NQ; q++)
                        atomicAdd(&tval[slot][v * NQ + q], (unsigned long long)fix_rn_any((double)val[v][q] * g.fix));
            }
        }
        if (__ballot(on && slot < 0)) {
            if (on && slot < 0) {
                const int bx = D == 2 ? box % g.nb : box, by = D == 2 ? box / g.nb : 0;
                for (int v = 0; v < NV; v++)
                    for (int q = 0; q < NQ; q++) {
                        const long node =
                            D == 2 ? (long)(bx * P + v / P) * g.n1d + by * P + v % P : (long)bx * P + v;
                        atomicAdd(&grid[node * 4 + q], (unsigned long long)fix_rn_any((double)val[v][q] * g.fix));
                    }
            }
        }
    };
    for (int chunk = 0; chunk < pts / (2 * kSpreadThreads); chunk++) {
        const long i = (long)bid * pts + (long)chunk * 2 * kSpreadThreads + 2 * (long)t;  // point a; b = i + 1
        float ya[D], yb[D];
        const bool oka = i < N, okb = i + 1 < N;
        if (okb) {
            if constexpr (D == 2) {
                const float4 v = reinterpret_cast<const float4 *>(Y)[i >> 1];
                ya[0] = v.x;
                ya[1] = v.y;
                yb[0] = v.z;
                yb[1] = v.w;
            } else {
                const float2 v = reinterpret_cast<const float2 *>(Y)[i >> 1];
                ya[0] = v.x;
                yb[0] = v.y;
            }
        } else if (oka) {
            load_pt<float, D>(Y, i, ya);
        }
        if (ctr) {  // pending centring (uniform branch)
#pragma unroll
            for (int d = 0; d < D; d++) {
                ya[d] = oka ? centred(ya[d], ctr[d]) : 0.f;
                yb[d] = okb ? centred(yb[d], ctr[d]) : 0.f;
            }
        }
        float val[NV][NQ];
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int q = 0; q < NQ; q++) val[v][q] = 0.f;
        // boxes (lanes past N: distinct negative keys, never a real box or run)
        int ka = -1 - 2 * lane, kb = -2 - 2 * lane;
        if (oka) ka = point_charges<P, D>(ya, g, val);
        if (okb) {
            double tb;
            kb = box_coord<D>((double)yb[0], g, tb);
            if constexpr (D == 2) kb += box_coord<D>((double)yb[1], g, tb) * g.nb;
        }
        const int kb_prev = __shfl_up(kb, 1, kWave), ka_next = __shfl_down(ka, 1, kWave),
                  kb_next = __shfl_down(kb, 1, kWave);
        const bool pure = ka == kb;                                          // a is in b's box
        const bool give = !pure && lane > 0 && ka == kb_prev;                // a continues the left lane's box
        const bool take = lane < kWave - 1 && ka_next == kb && ka_next != kb_next;  // the right lane's a is ours
        if (__ballot(oka && !pure && !give)) flush(oka && !pure && !give, ka, val);
        // a's charges: kept (pure), moved one lane to the left (give / take), or already flushed
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const float right = __int_as_float(
                    __builtin_amdgcn_update_dpp(0, __float_as_int(val[v][q]), 0x130, 0xf, 0xf, false));  // wave_shl:1
                const float own = pure ? val[v][q] : 0.f;
                val[v][q] = take ? own + right : own;
            }
        if (okb) {
            float vb[NV][NQ];
            (void)point_charges<P, D>(yb, g, vb);
#pragma unroll
            for (int v = 0; v < NV; v++)
#pragma unroll
                for (int q = 0; q < NQ; q++) val[v][q] += vb[v][q];
        }
        // runs of equal b boxes among consecutive lanes: a segmented inclusive scan over the wave (DPP row
        // shifts and row broadcasts, in the VALU) leaves each run's sum in its last lane.  The wave's
        // points and their order are fixed, so the float sums are bitwise repeatable.
        const bool head = lane == 0 || kb != kb_prev;
        const unsigned long long heads = __ballot(head);
        const int run = __popcll(heads & ((2ull << lane) - 1ull));  // 1-based run of this lane
        seg_scan_step<0x111, 0xf, NV, NQ>(val, run);  // row_shr:1
        seg_scan_step<0x112, 0xf, NV, NQ>(val, run);  // row_shr:2
        seg_scan_step<0x114, 0xf, NV, NQ>(val, run);  // row_shr:4
        seg_scan_step<0x118, 0xf, NV, NQ>(val, run);  // row_shr:8
        seg_scan_step<0x142, 0xa, NV, NQ>(val, run);  // row_bcast:15 into rows 1 and 3
        seg_scan_step<0x143, 0xc, NV, NQ>(val, run);  // row_bcast:31 into rows 2 and 3
        const bool tail = lane == kWave - 1 || kb != kb_next;
        flush(tail && kb >= 0, kb, val);
    }
    __syncthreads();
    // table -> grid
    for (int o = t; o < NS * NV * NQ; o += kSpreadThreads) {
        const int slot = o / (NV * NQ), r = o - slot * (NV * NQ);
        const int box = tkey[slot];
        if (box < 0) continue;
        const unsigned long long sum = tval[slot][r];
        if (sum == 0ull) continue;
        const int v = r / NQ, q = r - v * NQ;
        const int bx = D == 2 ? box % g.nb : box, by = D == 2 ? box / g.nb : 0;
        const long node = D == 2 ? (long)(bx * P + v / P) * g.n1d + by * P + v % P : (long)bx * P + v;
        atomicAdd(&grid[node * 4 + q], sum);
    }
}

template <int P, int D>
__global__ __launch_bounds__(kSpreadThreads) __attribute__((amdgpu_waves_per_eu(D == 2 && P <= 3 ? 4 : 1))) void k_spread_f32(const float *__restrict__ Y, int N, int pts,
                                                               const DevGeo *__restrict__ dg,
                                                               unsigned long long *__restrict__ grid,
                                                               const double *__restrict__ ctr) {
    __shared__ SpreadLds<P, D> L;
    if (!dg->live) return;
    spread_block<P, D>(Y, N, pts, dg, grid, blockIdx.x, L, ctr);
}

// The same with a host-planned geometry record passed by value; the first workgroup also stores it to
// `out` for the passes that follow on the stream (this replaces k_set_geo, which, queued behind the
// attractive kernel's workgroups, would hold the spread back).
template <int P, int D>
__global__ __launch_bounds__(kSpreadThreads) __attribute__((amdgpu_waves_per_eu(D == 2 && P <= 3 ? 4 : 1))) void k_spread_f32_v(const float *__restrict__ Y, int N, int pts,
                                                               const DevGeo v, DevGeo *__restrict__ out,
                                                               unsigned long long *__restrict__ grid,
                                                               const double *__restrict__ ctr) {
    __shared__ SpreadLds<P, D> L;
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = v;
    if (!v.live) return;
    spread_block<P, D>(Y, N, pts, &v, grid, blockIdx.x, L, ctr);
}

// int64 grid -> float4 (A, c, b_x, b_y) and clear, for the 1D path (the 2D row FFT does this itself).
__global__ __launch_bounds__(kThreads) void k_grid_to_f32(const DevGeo *__restrict__ dg,
                                                          unsigned long long *__restrict__ grid,
                                                          float4 *__restrict__ out) {
    if (!dg->live) return;
    const Geo &g = dg->g;
    const double inv = 1.0 / g.fix;
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < g.n1d; n += (long)gridDim.x * kThreads) {
        long long s[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            s[q] = (long long)grid[n * 4 + q];
            grid[n * 4 + q] = 0ull;
        }
        out[n] = make_float4((float)(s[0] * inv), (float)(s[1] * inv * g.bw0 * g.bw0), (float)(s[2] * inv * g.bw0),
                             (float)(s[3] * inv * g.bw0));
    }
}

// ---------------------------------------------------------------- 2D FFT passes
// Every pass reads the step's geometry and plan from DevGeo and walks its lines (rows u in [0, n1d) or
// column frequencies k in [0, H)) over XCD slots with a workgroup stride, so one launch sized for the
// largest grid serves any grid of the step; each line ends with a barrier because the next one reuses
// the LDS.  Twiddles are staged in LDS once per workgroup, behind the lines.

// Kernel spectra, row pass: for lattice row a (x offset a h), the even real lines K_q(a, |b|) and
// K_1(a, |b|) over b in (-n1d, n1d) wrapped into Mp, packed as one complex line; its DFT is
// T_q + i T_1 (both real).  Trow[k][a] = (T_q(a, k), T_1(a, k)), k in [0, H).  Runs only when the
// step marks the spectra stale.
template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void kern_rows_line(const DevGeo *__restrict__ dg, cx<T> *buf, const cx<T> *twl,
                                               int kind_q, int kind_1, double df, cx<T> *__restrict__ Trow, int a,
                                               int nl = 1) {
    // rows a .. a + nl - 1 (nl Mp <= MAXTOT), line r of the buffer for row a + r
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    for (int b = threadIdx.x; b < nl * Mp; b += NTHR) buf[lp<PAD>(b)] = cx<T>{(T)0, (T)0};
    __syncthreads();
    for (int q = threadIdx.x; q < nl * n1d; q += NTHR) {
        const int r = q / n1d, b = q - r * n1d;
        const double du = (a + r) * g.h, dv = b * g.h;
        const double r2 = du * du + dv * dv;
        cx<T> val{kernel_value<T>(kind_q, r2, df), kernel_value<T>(kind_1, r2, df)};
        buf[lp<PAD>(r * Mp + b)] = val;
        if (b > 0) buf[lp<PAD>(r * Mp + Mp - b)] = val;
    }
    __syncthreads();
    fft_lines<T, NTHR, MAXTOT, PAD>(buf, dg->pl, nl, Mp, twl, false);
    for (int q = threadIdx.x; q < nl * H; q += NTHR) {  // the rows of one frequency are contiguous in Trow
        const int k = q / nl, r = q - k * nl;
        Trow[(long)k * n1d + a + r] = buf[lp<PAD>(r * Mp + k)];
    }
    __syncthreads();
}

template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void kern_rows_body(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                               int kind_q, int kind_1, double df, cx<T> *__restrict__ Trow) {
    if (!dg->live || !dg->kspec_stale) return;
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d;
    const cx<T> *twl = stage_twiddles(buf + lspan<PAD>(Mp), tw, Mp);
    const int npad = 8 * ((n1d + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int a = xcd_slot_row(slot, n1d);
        if (a >= n1d) continue;
        kern_rows_line<T, NTHR, MAXTOT, PAD>(dg, buf, twl, kind_q, kind_1, df, Trow, a);
    }
}

template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_kern_rows(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                    int kind_q, int kind_1, double df, cx<T> *__restrict__ Trow) {
    kern_rows_body<T, NTHR, MAXTOT, PAD>(dg, tw, kind_q, kind_1, df, Trow);
}

// The same with a host-planned geometry record passed by value, so that the pass need not wait for the
// record to reach device memory (it runs on another queue than the spread that stores it).
template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_kern_rows_v(const DevGeo v, const cx<T> *__restrict__ tw, int kind_q,
                                                      int kind_1, double df, cx<T> *__restrict__ Trow) {
    kern_rows_body<T, NTHR, MAXTOT, PAD>(&v, tw, kind_q, kind_1, df, Trow);
}

// buf[q] = f(q) for q < total, the loads of a thread issued eight at a time before their LDS writes, so
// a workgroup waits for memory once per eight elements per thread instead of once per element.
template <int NTHR, bool PAD, typename V, typename F>
__device__ __forceinline__ void lds_fill(V *buf, int total, F &&f) {
    constexpr int B = 8;
    for (int q0 = threadIdx.x; q0 < total; q0 += B * NTHR) {
        V v[B];
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int q = q0 + s * NTHR;
            if (q < total) v[s] = f(q);
        }
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int q = q0 + s * NTHR;
            if (q < total) buf[lp<PAD>(q)] = v[s];
        }
    }
}

// lds_fill with element q stored at slot(q) instead of q (a permutation of [0, total)), so that the loads
// of consecutive threads can follow the source's order rather than the buffer's.
template <int NTHR, bool PAD, typename V, typename S, typename F>
__device__ __forceinline__ void lds_fill_slot(V *buf, int total, S &&slot, F &&f) {
    constexpr int B = 8;
    for (int q0 = threadIdx.x; q0 < total; q0 += B * NTHR) {
        V v[B];
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int q = q0 + s * NTHR;
            if (q < total) v[s] = f(q);
        }
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int q = q0 + s * NTHR;
            if (q < total) buf[lp<PAD>(slot(q))] = v[s];
        }
    }
}

// Kernel spectra, column pass: for column frequency k, the line z(a) = T_q(|a|,k)(1 + a h) + i T_1(|a|,k)
// (a signed).  Its DFT is C_q(m,k) + i (C_1(m,k) - S(m,k)) with C even and S odd in m, so
//   C_q = Re, C_1 = (Im(m) + Im(-m)) / 2, S = (Im(-m) - Im(m)) / 2,
// where S(m,k) = sum_{a,b} K_q(a,b) a h sin(2 pi m a / Mp) cos(2 pi k b / Mp): the DFT of the odd
// kernel K_q(a,b) a h is -i S(m,k) and that of K_q(a,b) b h is -i S(k,m) (K_q symmetric in a, b).
// Stored for m, k in [0, H):  KQ[k][m] = C_q(m,k), K1[k][m] = C_1(m,k), KS[k][m] = S(m,k),
// KT[k][m] = S(k,m) (written by column m's workgroup).
template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void kern_cols_line(const DevGeo *__restrict__ dg, cx<T> *buf, const cx<T> *twl,
                                               const cx<T> *__restrict__ Trow, int need_s, T *__restrict__ KQ,
                                               T *__restrict__ K1, T *__restrict__ KS, T *__restrict__ KT, int k,
                                               int nc = 1) {
    // columns k .. k + nc - 1 (nc Mp <= MAXTOT), line c of the buffer for column k + c
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    lds_fill<NTHR, PAD>(buf, nc * Mp, [&](int q) {
        const int c = q / Mp, a = q - c * Mp;
        int as;
        if (a < n1d) as = a;
        else if (a > Mp - n1d) as = a - Mp;
        else return cx<T>{(T)0, (T)0};
        const cx<T> t = Trow[(long)(k + c) * n1d + (as < 0 ? -as : as)];
        const T w = need_s ? (T)1 + (T)(as * g.h) : (T)1;
        return cx<T>{t.x * w, t.y};
    });
    __syncthreads();
    fft_lines<T, NTHR, MAXTOT, PAD>(buf, dg->pl, nc, Mp, twl, false);
    for (int q = threadIdx.x; q < nc * H; q += NTHR) {
        const int c = q / H, m = q - c * H, kc = k + c;
        const cx<T> z = buf[lp<PAD>(c * Mp + m)];
        const T im_neg = buf[lp<PAD>(c * Mp + (m == 0 ? 0 : Mp - m))].y;
        KQ[(long)kc * H + m] = z.x;
        K1[(long)kc * H + m] = (z.y + im_neg) * (T)0.5;
        if (need_s) {
            const T sv = (im_neg - z.y) * (T)0.5;
            KS[(long)kc * H + m] = sv;
            KT[(long)m * H + kc] = sv;
        }
    }
    __syncthreads();
}

template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void kern_cols_body(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                               const cx<T> *__restrict__ Trow, int need_s, T *__restrict__ KQ,
                                               T *__restrict__ K1, T *__restrict__ KS, T *__restrict__ KT) {
    if (!dg->live || !dg->kspec_stale) return;
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, H = g.H;
    const cx<T> *twl = stage_twiddles(buf + lspan<PAD>(Mp), tw, Mp);
    const int npad = 8 * ((H + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int k = xcd_slot_row(slot, H);
        if (k >= H) continue;
        kern_cols_line<T, NTHR, MAXTOT, PAD>(dg, buf, twl, Trow, need_s, KQ, K1, KS, KT, k);
    }
}

template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_kern_cols(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                    const cx<T> *__restrict__ Trow, int need_s,
                                                    T *__restrict__ KQ, T *__restrict__ K1, T *__restrict__ KS,
                                                    T *__restrict__ KT) {
    kern_cols_body<T, NTHR, MAXTOT, PAD>(dg, tw, Trow, need_s, KQ, K1, KS, KT);
}

template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_kern_cols_v(const DevGeo v, const cx<T> *__restrict__ tw,
                                                      const cx<T> *__restrict__ Trow, int need_s,
                                                      T *__restrict__ KQ, T *__restrict__ K1, T *__restrict__ KS,
                                                      T *__restrict__ KT) {
    kern_cols_body<T, NTHR, MAXTOT, PAD>(&v, tw, Trow, need_s, KQ, K1, KS, KT);
}

// Row FFT, forward.  grid: [n1d][n1d][sin] real (u = x node, v = y node); terms 2l, 2l+1 packed as
// one complex line.  Output spec[t][k][u] = DFT over v of term t of row u, k in [0, H) (two real
// spectra unpacked from the packed line: A = (Z(k) + conj Z(-k)) / 2, B = -i (Z(k) - conj Z(-k)) / 2).
// With fixgrid != null the input is the exact int64 grid of k_spread_f32 ([u][v][4] = A, c, b_x, b_y in
// 2^-38 units, offsets in box widths); it is converted and cleared here (each row read once).
template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void rows_fwd_line(const DevGeo *__restrict__ dg, cx<T> *buf, const cx<T> *twl,
                                              const T *__restrict__ grid, int sin, int nterms,
                                              cx<T> *__restrict__ spec, unsigned long long *__restrict__ fixgrid,
                                              int u0, int nl = 1) {
    // rows u0 .. u0 + nl - 1 (nl lines Mp <= MAXTOT); buffer line r lines + l for row u0 + r
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    const int lines = (nterms + 1) / 2;
    const double inv = 1.0 / g.fix;
    const double unit[4] = {inv, inv * g.bw0 * g.bw0, inv * g.bw0, inv * g.bw0};
    const long plane = (long)H * n1d;
    lds_fill<NTHR, PAD>(buf, nl * lines * Mp, [&](int q) {
        const int ln = q / Mp, v = q - ln * Mp;
        const int r = ln / lines, l = ln - r * lines, u = u0 + r;
        cx<T> z{(T)0, (T)0};
        if (v < n1d) {
            if (fixgrid) {
                unsigned long long *src = fixgrid + ((long)u * n1d + v) * 4 + 2 * l;
                z.x = (T)((double)(long long)src[0] * unit[2 * l]);
                z.y = (T)((double)(long long)src[1] * unit[2 * l + 1]);
                src[0] = 0ull;
                src[1] = 0ull;
            } else {
                const T *src = grid + ((long)u * n1d + v) * sin + 2 * l;
                z.x = src[0];
                if (2 * l + 1 < nterms) z.y = src[1];
            }
        }
        return z;
    });
    __syncthreads();
    fft_lines<T, NTHR, MAXTOT, PAD>(buf, dg->pl, nl * lines, Mp, twl, false);
    for (int q = threadIdx.x; q < nl * lines * H; q += NTHR) {  // rows fastest: contiguous in spec
        const int lk = q / nl, r = q - lk * nl;
        const int l = lk / H, k = lk - l * H, b = (r * lines + l) * Mp, u = u0 + r;
        const cx<T> Z = buf[lp<PAD>(b + k)];
        const cx<T> Zc = conj(buf[lp<PAD>(b + (k == 0 ? 0 : Mp - k))]);
        spec[(2 * l) * plane + (long)k * n1d + u] = (T)0.5 * (Z + Zc);
        if (2 * l + 1 < nterms) spec[(2 * l + 1) * plane + (long)k * n1d + u] = (T)0.5 * mul_negi(Z - Zc);
    }
    __syncthreads();
}

template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_rows_fwd(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                   const T *__restrict__ grid, int sin, int nterms,
                                                   cx<T> *__restrict__ spec,
                                                   unsigned long long *__restrict__ fixgrid) {
    if (!dg->live) return;
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d;
    const int lines = (nterms + 1) / 2;
    const cx<T> *twl = stage_twiddles(buf + lspan<PAD>(lines * Mp), tw, Mp);
    const int npad = 8 * ((n1d + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int u = xcd_slot_row(slot, n1d);
        if (u >= n1d) continue;
        rows_fwd_line<T, NTHR, MAXTOT, PAD>(dg, buf, twl, grid, sin, nterms, spec, fixgrid, u);
    }
}

// Column pass, FP32 form.  For column frequency k: DFT over u of A, c, b_x, b_y; then with K_q, K_1
// (even) and the odd kernels K_x = -i S(m,k), K_y = -i S(k,m):
//   V1 = K_q A,  G_x = K_x A - K_q b_x,  G_y = K_y A - K_q b_y    (inverse-transformed over u)
// and the Parseval sums of Z (k weighted 1 at k = 0, Mp/2 and 2 elsewhere):
//   z0 = sum K_1 |A|^2, z1 = sum Re(conj(c) K_q A), z2 = sum Re(conj(b).K_D A), z3 = sum K_q |b|^2.
// Lanes of the column pass's Z sums (the single-context pass's workgroup size).
constexpr int kColZLanes = 512;

// Sums over NTHR * VPT virtual lanes (lane t + j NTHR holds v[i][j] of thread t) of NV values at once, each
// in the order block_sum takes with that many threads: each wavefront of lanes by shuffles, then the
// wavefront totals in lane order from 0.0 (results valid in thread 0).  sh: NV VPT NTHR / kWave doubles.
template <int NTHR, int VPT, int NV>
__device__ __forceinline__ void lanes_sum(const double (&v)[NV][VPT], double wk, double *sh, double (&out)[NV]) {
    constexpr int NW = NTHR / kWave;
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    double t[NV][VPT];
#pragma unroll
    for (int i = 0; i < NV; i++)
#pragma unroll
        for (int j = 0; j < VPT; j++) t[i][j] = wave_sum(v[i][j] * wk);
    __syncthreads();
    if (l == 0)
#pragma unroll
        for (int i = 0; i < NV; i++)
#pragma unroll
            for (int j = 0; j < VPT; j++) sh[(i * VPT + j) * NW + w] = t[i][j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double r = 0.0;
        if (threadIdx.x == 0)
            for (int q = 0; q < VPT * NW; q++) r += sh[i * VPT * NW + q];
        out[i] = r;
    }
}

// Column frequencies k .. k + nc - 1 (4 nc Mp <= MAXTOT) of the column pass: 4 lines per column forward
// (A, c, b_x, b_y), the kernel spectra products into 3 lines per column (written over the inputs, column
// after column, so that the products of column c sit in lines 3c .. 3c + 2), Parseval's Z terms, and the
// 3 nc lines inverse.
template <int NTHR, int MAXTOT>
__device__ __forceinline__ void cols_f32_line(const DevGeo *__restrict__ dg, cx<float> *buf, const cx<float> *twl,
                                              const cx<float> *__restrict__ spec, const float *__restrict__ KQ,
                                              const float *__restrict__ K1, const float *__restrict__ KS,
                                              const float *__restrict__ KT, cx<float> *__restrict__ ospec,
                                              double *__restrict__ zpart, double *shz, int k0, int nc = 1) {
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    const long plane = (long)H * n1d;
    lds_fill<NTHR, true>(buf, 4 * nc * Mp, [&](int q) {
        const int ln = q / Mp, u = q - ln * Mp, c = ln >> 2, t = ln & 3;
        return u < n1d ? spec[t * plane + (long)(k0 + c) * n1d + u] : cx<float>{0.f, 0.f};
    });
    __syncthreads();
    fft_lines<float, NTHR, MAXTOT, true>(buf, dg->pl, 4 * nc, Mp, twl, false);
    // Z partial sums in the order of kColZLanes lanes whatever the workgroup size: thread t stands for
    // lanes t + j NTHR, each summing m = lane, lane + kColZLanes, ...
    constexpr int VPT = kColZLanes / NTHR;
    static_assert(VPT * NTHR == kColZLanes && NTHR % kWave == 0, "column pass: threads must divide the Z lanes");
    for (int c = 0; c < nc; c++) {
        const int k = k0 + c;
        const double wk = (k == 0 || 2 * k == Mp) ? 1.0 : 2.0;
        const int in = 4 * c * Mp, outl = 3 * c * Mp;
        double z[4][VPT];
#pragma unroll
        for (int j = 0; j < VPT; j++) z[0][j] = z[1][j] = z[2][j] = z[3][j] = 0.0;
#pragma unroll
        for (int j = 0; j < VPT; j++)
        for (int m = threadIdx.x + j * NTHR; m < Mp; m += kColZLanes) {
            const int mb = m <= Mp / 2 ? m : Mp - m;
            const float sg = m <= Mp / 2 ? 1.f : -1.f;
            const float cq = KQ[(long)k * H + mb], c1 = K1[(long)k * H + mb];
            const float sx = sg * KS[(long)k * H + mb], sy = KT[(long)k * H + mb];
            const cx<float> A = buf[lp<true>(in + m)], C = buf[lp<true>(in + Mp + m)],
                            BX = buf[lp<true>(in + 2 * Mp + m)], BY = buf[lp<true>(in + 3 * Mp + m)];
            const cx<float> KxA = mul_negi(sx * A), KyA = mul_negi(sy * A);
            buf[lp<true>(outl + m)] = cq * A;
            buf[lp<true>(outl + Mp + m)] = KxA - cq * BX;
            buf[lp<true>(outl + 2 * Mp + m)] = KyA - cq * BY;
            const double a2 = (double)A.x * A.x + (double)A.y * A.y;
            z[0][j] += (double)c1 * a2;
            z[1][j] += (double)cq * ((double)C.x * A.x + (double)C.y * A.y);
            // Re(conj(B) (-i s) A) = s Im(conj(B) A)
            z[2][j] += (double)sx * ((double)BX.x * A.y - (double)BX.y * A.x) +
                       (double)sy * ((double)BY.x * A.y - (double)BY.y * A.x);
            z[3][j] += (double)cq * ((double)BX.x * BX.x + (double)BX.y * BX.y + (double)BY.x * BY.x +
                                     (double)BY.y * BY.y);
        }
        double zs[4];
        lanes_sum<NTHR, VPT, 4>(z, wk, shz, zs);  // its barriers also order this column's writes before the next
        if (threadIdx.x == 0)
#pragma unroll
            for (int i = 0; i < 4; i++) zpart[k * 4 + i] = zs[i];
    }
    __syncthreads();
    fft_lines<float, NTHR, MAXTOT, true>(buf, dg->pl, 3 * nc, Mp, twl, true);
    for (int q = threadIdx.x; q < 3 * nc * n1d; q += NTHR) {
        const int ln = q / n1d, u = q - ln * n1d, c = ln / 3, t = ln - 3 * c;
        ospec[t * plane + (long)(k0 + c) * n1d + u] = buf[lp<true>(ln * Mp + u)];
    }
    __syncthreads();
}

template <int NTHR, int MAXTOT>
__global__ __launch_bounds__(NTHR) void k_cols_f32(const DevGeo *__restrict__ dg, const cx<float> *__restrict__ tw,
                                                   const cx<float> *__restrict__ spec, const float *__restrict__ KQ,
                                                   const float *__restrict__ K1, const float *__restrict__ KS,
                                                   const float *__restrict__ KT, cx<float> *__restrict__ ospec,
                                                   double *__restrict__ zpart) {
    if (!dg->live) return;
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    __shared__ double shz[4 * kColZLanes / kWave];
    const int Mp = g.Mp, H = g.H;
    const cx<float> *twl = stage_twiddles(buf + lspan<true>(4 * Mp), tw, Mp);
    const int npad = 8 * ((H + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int k = xcd_slot_row(slot, H);
        if (k >= H) continue;
        cols_f32_line<NTHR, MAXTOT>(dg, buf, twl, spec, KQ, K1, KS, KT, ospec, zpart, shz, k);
    }
}

// Column pass, plain (FP64).  Each term: DFT over u, times the real even kernel spectrum C(m,k),
// inverse DFT over u.
template <typename T, int NTHR, int MAXTOT>
__global__ __launch_bounds__(NTHR) void k_cols_plain(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                     const cx<T> *__restrict__ spec, int nterms,
                                                     const T *__restrict__ KC, cx<T> *__restrict__ ospec) {
    if (!dg->live) return;
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    const long plane = (long)H * n1d;
    const int lines_max = MAXTOT / Mp;
    const cx<T> *twl = stage_twiddles(buf + min(nterms, lines_max) * Mp, tw, Mp);
    const int npad = 8 * ((H + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int k = xcd_slot_row(slot, H);
        if (k >= H) continue;
        for (int t0 = 0; t0 < nterms; t0 += lines_max) {
            const int lines = min(lines_max, nterms - t0);
            for (int q = threadIdx.x; q < lines * Mp; q += NTHR) {
                const int t = q / Mp, u = q - t * Mp;
                buf[q] = u < n1d ? spec[(t0 + t) * plane + (long)k * n1d + u] : cx<T>{(T)0, (T)0};
            }
            __syncthreads();
            fft_lines<T, NTHR, MAXTOT>(buf, dg->pl, lines, Mp, twl, false);
            for (int q = threadIdx.x; q < lines * Mp; q += NTHR) {
                const int t = q / Mp, m = q - t * Mp;
                const int mb = m <= Mp / 2 ? m : Mp - m;
                buf[q] = KC[(long)k * H + mb] * buf[q];
            }
            __syncthreads();
            fft_lines<T, NTHR, MAXTOT>(buf, dg->pl, lines, Mp, twl, true);
            for (int q = threadIdx.x; q < lines * n1d; q += NTHR) {
                const int t = q / n1d, u = q - t * n1d;
                ospec[(t0 + t) * plane + (long)k * n1d + u] = buf[t * Mp + u];
            }
            __syncthreads();
        }
    }
}

// Row FFT, inverse.  ospec[t][k][u], k in [0, H): half spectra over v of real rows.  Terms 2l, 2l+1 are
// rebuilt into one Hermitian-extended complex line (self-conjugate bins taken real), inverse-
// transformed, scaled by 1 / Mp^2, and written to out[u][v][sout + ...].
template <typename T, int NTHR, int MAXTOT, bool PAD>
__device__ __forceinline__ void rows_inv_line(const DevGeo *__restrict__ dg, cx<T> *buf, const cx<T> *twl,
                                              const cx<T> *__restrict__ ospec, int nterms, T *__restrict__ out,
                                              int sout, int u0, int nl = 1) {
    // rows u0 .. u0 + nl - 1 (nl lines Mp <= MAXTOT); buffer line r lines + l for row u0 + r.  Element q
    // of the fill takes rows fastest, so that the reads of consecutive threads are contiguous in ospec.
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d, H = g.H;
    const T scale = (T)(1.0 / ((double)Mp * Mp));
    const long plane = (long)H * n1d;
    const int lines = (nterms + 1) / 2;
    auto slot = [&](int q) {
        const int lk = q / nl, r = q - lk * nl, l = lk / Mp, k = lk - l * Mp;
        return (r * lines + l) * Mp + k;
    };
    lds_fill_slot<NTHR, PAD>(buf, nl * lines * Mp, slot, [&](int q) {
        const int lk = q / nl, r = q - lk * nl, l = lk / Mp, k = lk - l * Mp, u = u0 + r;
        const bool lo = k < H;
        const int kk = lo ? k : Mp - k;
        cx<T> f1 = ospec[(2 * l) * plane + (long)kk * n1d + u];
        cx<T> f2 = 2 * l + 1 < nterms ? ospec[(2 * l + 1) * plane + (long)kk * n1d + u] : cx<T>{(T)0, (T)0};
        if (kk == 0 || 2 * kk == Mp) {
            f1.y = (T)0;
            f2.y = (T)0;
        }
        if (!lo) {
            f1 = conj(f1);
            f2 = conj(f2);
        }
        return f1 + mul_i(f2);
    });
    __syncthreads();
    fft_lines<T, NTHR, MAXTOT, PAD>(buf, dg->pl, nl * lines, Mp, twl, true);
    for (int q = threadIdx.x; q < nl * lines * n1d; q += NTHR) {
        const int ln = q / n1d, v = q - ln * n1d, r = ln / lines, l = ln - r * lines;
        const cx<T> z = buf[lp<PAD>(ln * Mp + v)];
        T *dst = out + ((long)(u0 + r) * n1d + v) * sout + 2 * l;
        dst[0] = z.x * scale;
        if (2 * l + 1 < nterms) dst[1] = z.y * scale;
    }
    __syncthreads();
}

// zpart != null: the first workgroup also computes the FP32 form's final Z (k_zfinal's reduction and
// order; NTHR == kThreads) before its lines, one launch fewer on the step's critical path.
template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__global__ __launch_bounds__(NTHR) void k_rows_inv(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                   const cx<T> *__restrict__ ospec, int nterms,
                                                   T *__restrict__ out, int sout, const double *__restrict__ zpart,
                                                   int df_one, double Nz, double *__restrict__ zout) {
    if (!dg->live) return;
    if (zpart && blockIdx.x == 0) {
        __shared__ double shz[NTHR / kWave];
        zfinal_block(dg, zpart, df_one, Nz, zout, shz);
    }
    const Geo &g = dg->g;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d;
    const int lines = (nterms + 1) / 2;
    const cx<T> *twl = stage_twiddles(buf + lspan<PAD>(lines * Mp), tw, Mp);
    const int npad = 8 * ((n1d + 7) / 8);
    for (int slot = blockIdx.x; slot < npad; slot += gridDim.x) {
        const int u = xcd_slot_row(slot, n1d);
        if (u >= n1d) continue;
        rows_inv_line<T, NTHR, MAXTOT, PAD>(dg, buf, twl, ospec, nterms, out, sout, u);
    }
}

// ---------------------------------------------------------------- 1D convolution, one workgroup
// The 1D grid is at most a few thousand nodes: one block of NTHR threads does the whole convolution.
// The kernel line is transformed first and its spectra kept in registers (thread t owns bins
// m = t + c NTHR); LDS is then reused for the data lines.  Requires Mp <= MAXM.
// FP32 form (f32form = 1): kernel line K_q(|a|)(1 + a h) + i K_1(|a|) gives C_q, C_1 and S as in
// k_kern_cols; data lines z = A + i b and c; V1 = C_q A, G = -i S A - C_q b, one inverse of V1 + i G;
// Z partial sums z0..z3 as in k_cols_f32 (full spectrum, no half-plane weights).
// Plain (f32form = 0): real even spectrum of K_q (or K_1 when use_k1) times each packed pair of terms.
// The kernel line of k_conv1d: the kernel values at the grid offsets, circularly embedded, and their
// transform, left in buf[0, Mp).  The arithmetic of every element is the same for any NTHR and MAXM.
template <typename T, int NTHR, int MAXM>
__device__ __forceinline__ void conv1d_kernel_fft(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                  int kind_q, int kind_1, double df, int f32form, int use_k1,
                                                  cx<T> *buf) {
    const Geo &g = dg->g;
    const int Mp = g.Mp, n1d = g.n1d;
    for (int q = threadIdx.x; q < Mp; q += NTHR) buf[q] = cx<T>{(T)0, (T)0};
    __syncthreads();
    for (int a = threadIdx.x; a < n1d; a += NTHR) {
        const double d = a * g.h;
        const T kq = kernel_value<T>(kind_q, d * d, df), k1 = kernel_value<T>(kind_1, d * d, df);
        if (f32form) {
            buf[a] = cx<T>{kq * ((T)1 + (T)d), k1};
            if (a > 0) buf[Mp - a] = cx<T>{kq * ((T)1 - (T)d), k1};
        } else {
            buf[a] = cx<T>{use_k1 ? k1 : kq, (T)0};
            if (a > 0) buf[Mp - a] = buf[a];
        }
    }
    __syncthreads();
    fft_lines<T, NTHR, MAXM>(buf, dg->pl, 1, Mp, tw, false);
}

// Bin m of the transformed kernel line: C_q (the squared kernel's, real), and C_1, S, the even and odd
// parts of the imaginary channel (the FP32 form carries the kernel and its d-weighted copy in one line).
template <typename T>
__device__ __forceinline__ void kernel_line_bin(const cx<T> *buf, int m, int Mp, T &cq, T &c1, T &ss) {
    const cx<T> kz = buf[m];
    const T imn = buf[m == 0 ? 0 : Mp - m].y;
    cq = kz.x;
    c1 = (kz.y + imn) * (T)0.5;
    ss = (imn - kz.y) * (T)0.5;
}

// The FP32 form's kernel line alone, into kline[3][Mp] = (C_q, C_1, S), with the step's geometry record
// by value: one workgroup on a stream of its own from the step's start, so that k_conv1d is left with the
// data lines.  Sized to share a CU with the 1D attractive kernel, so that it does not wait for that
// kernel's CUs: one line of LDS (32 KB at Mp = 4096, the 1D limit, beside its 128 KB of the CU's 160)
// and one wave per SIMD (NTHR = 256) at 80 VGPRs beside its four at 88 (of 512).
template <typename T, int NTHR, int MAXM>
__global__ __launch_bounds__(NTHR) void k_kline1d(const DevGeo v, const cx<T> *__restrict__ tw, int kind_q, int kind_1,
                                                  double df, T *__restrict__ kline) {
    if (!v.live) return;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    conv1d_kernel_fft<T, NTHR, MAXM>(&v, tw, kind_q, kind_1, df, 1, 0, buf);
    const int Mp = v.g.Mp;
    for (int m = threadIdx.x; m < Mp; m += NTHR) {
        T cq, c1, ss;
        kernel_line_bin(buf, m, Mp, cq, c1, ss);
        kline[m] = cq;
        kline[Mp + m] = c1;
        kline[2 * Mp + m] = ss;
    }
}

// kline (FP32 form, nullable): the kernel line's spectra from k_kline1d instead of computing them here.
template <typename T, int NTHR, int MAXM>
__global__ __launch_bounds__(NTHR) void k_conv1d(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw, int kind_q,
                                                 int kind_1, double df, int f32form, const T *__restrict__ grid,
                                                 int sin, int nterms, int use_k1, T *__restrict__ out, int sout,
                                                 double *__restrict__ zsum, const T *__restrict__ kspec,
                                                 const T *__restrict__ kline) {
    constexpr int NC = (MAXM + NTHR - 1) / NTHR;
    if (!dg->live) return;
    const Geo &g = dg->g;
    const FftPlan &pl = dg->pl;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d;
    T cq[NC], c1[NC], ss[NC];
    if (kspec) {  // plain path with a caller-given spectrum (mirror tier)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int m = threadIdx.x + c * NTHR;
            cq[c] = m < Mp ? kspec[m] : (T)0;
            c1[c] = ss[c] = (T)0;
        }
    } else if (kline) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int m = threadIdx.x + c * NTHR;
            cq[c] = m < Mp ? kline[m] : (T)0;
            c1[c] = m < Mp ? kline[Mp + m] : (T)0;
            ss[c] = m < Mp ? kline[2 * Mp + m] : (T)0;
        }
    } else {
        conv1d_kernel_fft<T, NTHR, MAXM>(dg, tw, kind_q, kind_1, df, f32form, use_k1, buf);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int m = threadIdx.x + c * NTHR;
            cq[c] = c1[c] = ss[c] = (T)0;
            if (m < Mp) kernel_line_bin(buf, m, Mp, cq[c], c1[c], ss[c]);
        }
        __syncthreads();
    }
    const T sc = (T)1 / (T)Mp;
    double z0 = 0, z1 = 0, z2 = 0, z3 = 0;
    if (f32form) {
        cx<T> *dl = buf;  // 2 lines: A + i b, c
        for (int q = threadIdx.x; q < 2 * Mp; q += NTHR) {
            const int l = q / Mp, a = q - l * Mp;
            cx<T> z{(T)0, (T)0};
            if (a < n1d) {
                const T *s = grid + (long)a * sin;
                z = l == 0 ? cx<T>{s[0], s[2]} : cx<T>{s[1], (T)0};
            }
            dl[q] = z;
        }
        __syncthreads();
        fft_lines<T, NTHR, MAXM>(dl, pl, 2, Mp, tw, false);
        cx<T> o[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int m = threadIdx.x + c * NTHR;
            o[c] = cx<T>{(T)0, (T)0};
            if (m < Mp) {
                const int mn = m == 0 ? 0 : Mp - m;
                const cx<T> Z = dl[m], Zc = conj(dl[mn]);
                const cx<T> A = (T)0.5 * (Z + Zc), B = (T)0.5 * mul_negi(Z - Zc);
                const cx<T> C = dl[Mp + m];
                const cx<T> V1 = cq[c] * A, G = mul_negi(ss[c] * A) - cq[c] * B;
                o[c] = V1 + mul_i(G);
                z0 += (double)c1[c] * ((double)A.x * A.x + (double)A.y * A.y);
                z1 += (double)cq[c] * ((double)C.x * A.x + (double)C.y * A.y);
                z2 += (double)ss[c] * ((double)B.x * A.y - (double)B.y * A.x);
                z3 += (double)cq[c] * ((double)B.x * B.x + (double)B.y * B.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int m = threadIdx.x + c * NTHR;
            if (m < Mp) dl[m] = o[c];
        }
        __syncthreads();
        fft_lines<T, NTHR, MAXM>(dl, pl, 1, Mp, tw, true);
        for (int a = threadIdx.x; a < n1d; a += NTHR) {
            out[(long)a * sout + 0] = dl[a].x * sc;
            out[(long)a * sout + 1] = dl[a].y * sc;
        }
    } else {
        const int dlines = (nterms + 1) / 2;
        for (int l = 0; l < dlines; l++) {
            for (int a = threadIdx.x; a < Mp; a += NTHR) {
                cx<T> z{(T)0, (T)0};
                if (a < n1d) {
                    const T *s = grid + (long)a * sin + 2 * l;
                    z = cx<T>{s[0], 2 * l + 1 < nterms ? s[1] : (T)0};
                }
                buf[a] = z;
            }
            __syncthreads();
            fft_lines<T, NTHR, MAXM>(buf, pl, 1, Mp, tw, false);
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int m = threadIdx.x + c * NTHR;
                if (m < Mp) buf[m] = cq[c] * buf[m];
            }
            __syncthreads();
            fft_lines<T, NTHR, MAXM>(buf, pl, 1, Mp, tw, true);
            for (int a = threadIdx.x; a < n1d; a += NTHR) {
                out[(long)a * sout + 2 * l] = buf[a].x * sc;
                if (2 * l + 1 < nterms) out[(long)a * sout + 2 * l + 1] = buf[a].y * sc;
            }
            __syncthreads();
        }
    }
    if (zsum) {
        __shared__ double shz[NTHR / kWave];
        z0 = block_sum(z0, shz);
        z1 = block_sum(z1, shz);
        z2 = block_sum(z2, shz);
        z3 = block_sum(z3, shz);
        if (threadIdx.x == 0) {
            zsum[0] = z0;
            zsum[1] = z1;
            zsum[2] = z2;
            zsum[3] = z3;
        }
    }
}

// Real even spectrum of one 1D kernel on the node lattice (precompute of nbodyfft.cpp:251-307,
// in this library's layout: C(m) for m in [0, Mp)).
template <typename T, int NTHR, int MAXM>
__global__ __launch_bounds__(NTHR) void k_kspec1d(const DevGeo *__restrict__ dg, const cx<T> *__restrict__ tw,
                                                  int kind, double df, T *__restrict__ out) {
    if (!dg->live) return;
    const Geo &g = dg->g;
    const FftPlan &pl = dg->pl;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int Mp = g.Mp, n1d = g.n1d;
    for (int q = threadIdx.x; q < Mp; q += NTHR) buf[q] = cx<T>{(T)0, (T)0};
    __syncthreads();
    for (int a = threadIdx.x; a < n1d; a += NTHR) {
        const double d = a * g.h;
        buf[a] = cx<T>{kernel_value<T>(kind, d * d, df), (T)0};
        if (a > 0) buf[Mp - a] = buf[a];
    }
    __syncthreads();
    fft_lines<T, NTHR, MAXM>(buf, pl, 1, Mp, tw, false);
    for (int m = threadIdx.x; m < Mp; m += NTHR) out[m] = buf[m].x;
}

// Final Z of the FP32 form from the per-column Parseval partials, in FP64 and a fixed order:
//   df = 1:  Z = (z0 + 2 z1 + 4 z2 - 2 z3) / Mp^D - N;   df != 1:  Z = z0 / Mp^D - N.
__device__ __forceinline__ void zfinal_block(const DevGeo *__restrict__ dg, const double *__restrict__ zpart,
                                             int df_one, double N, double *__restrict__ zout, double *sh) {
    const Geo &g = dg->g;
    const int ncols = g.D == 2 ? g.H : 1;
    const double norm = g.D == 2 ? 1.0 / ((double)g.Mp * g.Mp) : 1.0 / (double)g.Mp;
    double s = 0;
    for (int k = threadIdx.x; k < ncols; k += blockDim.x) {
        const double *z = zpart + 4 * k;
        s += df_one ? (z[0] + 2.0 * z[1] + 4.0 * z[2] - 2.0 * z[3]) : z[0];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double Z = s * norm - N;
        zout[0] = Z;
        zout[1] = 1.0 / Z;
    }
}

__global__ void k_zfinal(const DevGeo *__restrict__ dg, const double *__restrict__ zpart, int df_one, double N,
                         double *__restrict__ zout) {
    if (!dg->live) return;
    __shared__ double sh[kThreads / kWave];
    zfinal_block(dg, zpart, df_one, N, zout, sh);
}

// ---------------------------------------------------------------- interpolation + update, FP32 form
// Each thread takes kInterpPpt points, kThreads apart: points [bid kInterpPts, (bid + 1) kInterpPts) per
// workgroup.  All per-point loads of a thread are issued before any of them is used, then all grid
// reads, so that one wave keeps both points' memory traffic in flight.
constexpr int kInterpPpt = 2;
constexpr int kInterpPts = kInterpPpt * kThreads;

// Lagrange weights of one coordinate and the offsets from the nodes (in grid units times bw0).
template <int P, int D>
__device__ __forceinline__ int interp_axis(float yv, const Geo &g, float (&L)[P], float (&e)[P]) {
    double t;
    const int b = box_coord<D>((double)yv, g, t);
    float dt[P];
#pragma unroll
    for (int a = 0; a < P; a++) dt[a] = (float)(t - g.sp[a]);
#pragma unroll
    for (int a = 0; a < P; a++) {
        float v = 1.f;
#pragma unroll
        for (int c = 0; c < P; c++)
            if (c != a) v *= dt[c];
        L[a] = v * g.rdenf[a];
        e[a] = dt[a] * (float)g.bw0;
    }
    return b;
}

// Workgroup bid of the interpolation and update; the body of k_interp_f32 and of its batched form.
// sh: kThreads / kWave doubles of LDS (3 D kThreads / kWave with mmu).
template <int P, int D>
__device__ __forceinline__ void interp_block(float *__restrict__ Y, const float *__restrict__ pos,
                                             const float4 *__restrict__ gout, const DevGeo *__restrict__ dg,
                                             const double *__restrict__ zbuf, const StepDev *__restrict__ sd,
                                             float *__restrict__ dC, float *__restrict__ uY, float *__restrict__ gains,
                                             double *__restrict__ part, int N, int bid, double *sh,
                                             double *__restrict__ mmu = nullptr, long skip = -1,
                                             const double *__restrict__ ctr = nullptr) {
    constexpr int U = kInterpPpt;
    int idx[U];
    float y[U][D], pi[U][D], u[U][D], gn[U][D];
#pragma unroll
    for (int k = 0; k < U; k++) {
        idx[k] = bid * kInterpPts + k * kThreads + threadIdx.x;
        if (idx[k] < N) {  // the momentum state is there in every mode; it is used only by updates
            load_pt<float, D>(Y, idx[k], y[k]);
            if (ctr) {  // pending centring (uniform branch)
#pragma unroll
                for (int d = 0; d < D; d++) y[k][d] = centred(y[k][d], ctr[d]);
            }
            // read once per step: streamed past the caches
#pragma unroll
            for (int d = 0; d < D; d++) {
                pi[k][d] = __builtin_nontemporal_load(&pos[(long)idx[k] * D + d]);
                u[k][d] = __builtin_nontemporal_load(&uY[(long)idx[k] * D + d]);
                gn[k][d] = __builtin_nontemporal_load(&gains[(long)idx[k] * D + d]);
            }
        }
    }
    const Geo &g = dg->g;
    const StepDev up = *sd;  // mode 0: write dC only; 1: optimiser update of Y, uY, gains
    const float zinv = (float)zbuf[1];
    float Lx[U][P], ex[U][P], Ly[U][P], ey[U][P];
    float4 G[U][D == 2 ? P * P : P];
#pragma unroll
    for (int k = 0; k < U; k++) {
        if (idx[k] < N) {
            const int bx = interp_axis<P, D>(y[k][0], g, Lx[k], ex[k]);
            if constexpr (D == 2) {
                const int by = interp_axis<P, D>(y[k][1], g, Ly[k], ey[k]);
#pragma unroll
                for (int a = 0; a < P; a++) {
                    const float4 *rowp = gout + (long)(bx * P + a) * g.n1d + by * P;
#pragma unroll
                    for (int c = 0; c < P; c++) G[k][a * P + c] = rowp[c];
                }
            } else {
#pragma unroll
                for (int a = 0; a < P; a++) G[k][a] = gout[(long)bx * P + a];
            }
        }
    }
    double ysum[U][D];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int i = idx[k];
#pragma unroll
        for (int d = 0; d < D; d++) ysum[k][d] = 0.0;
        if (i >= N) continue;
        float neg[D];
#pragma unroll
        for (int d = 0; d < D; d++) neg[d] = 0.f;
        if constexpr (D == 2) {
#pragma unroll
            for (int a = 0; a < P; a++)
#pragma unroll
                for (int c = 0; c < P; c++) {
                    const float4 Gv = G[k][a * P + c];
                    const float w = Lx[k][a] * Ly[k][c];
                    neg[0] += w * (ex[k][a] * Gv.x + Gv.y);
                    neg[1] += w * (ey[k][c] * Gv.x + Gv.z);
                }
        } else {
#pragma unroll
            for (int a = 0; a < P; a++) neg[0] += Lx[k][a] * (ex[k][a] * G[k][a].x + G[k][a].y);
        }
        float dy[D];
#pragma unroll
        for (int d = 0; d < D; d++) dy[d] = pi[k][d] - neg[d] * zinv;
        if (up.mode == 0) {
#pragma unroll
            for (int d = 0; d < D; d++) dC[(long)i * D + d] = dy[d];
        } else {
            float yn[D];
            if (up.plain_descent) {
#pragma unroll
                for (int d = 0; d < D; d++) yn[d] = y[k][d] - dy[d];
            } else {
                float nrm = 0.f;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const float sdv = dy[d] == 0.f ? 0.f : (dy[d] < 0.f ? -1.f : 1.f);
                    const float su = u[k][d] == 0.f ? 0.f : (u[k][d] < 0.f ? -1.f : 1.f);
                    gn[k][d] = sdv != su ? gn[k][d] + 0.2f : gn[k][d] * 0.8f;
                    if (gn[k][d] < 0.01f) gn[k][d] = 0.01f;
                    u[k][d] = up.momentum * u[k][d] - up.lr * gn[k][d] * dy[d];
                    nrm += u[k][d] * u[k][d];
                }
                if (up.max_step_norm > 0.f) {
                    nrm = sqrtf(nrm);
                    if (nrm > up.max_step_norm) {
#pragma unroll
                        for (int d = 0; d < D; d++) u[k][d] *= up.max_step_norm / nrm;
                    }
                }
#pragma unroll
                for (int d = 0; d < D; d++) {
                    __builtin_nontemporal_store(u[k][d], &uY[(long)i * D + d]);
                    __builtin_nontemporal_store(gn[k][d], &gains[(long)i * D + d]);
                    yn[d] = y[k][d] + u[k][d];
                }
            }
#pragma unroll
            for (int d = 0; d < D; d++) {
                Y[(long)i * D + d] = yn[d];
                ysum[k][d] = (double)yn[d];
            }
        }
    }
    if (up.mode == 1 && !mmu) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            double t = ysum[0][d];
#pragma unroll
            for (int k = 1; k < U; k++) t += ysum[k][d];
            const double s = block_sum(t, sh);
            if (threadIdx.x == 0) part[bid * D + d] = s;
        }
    } else if (up.mode == 1) {
        // the same sums (block_sum's order), and with them in one pass the per coordinate max and min of the
        // updated Y (before its mean is taken off), the min without the flat element `skip`: k_mean_mm
        // turns them into those of the centred Y.  sh holds 3 D kThreads / kWave doubles.
        constexpr int NW = kThreads / kWave;
        double v[3 * D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            double t = ysum[0][d], mx = -INFINITY, mn = INFINITY;
#pragma unroll
            for (int k = 1; k < U; k++) t += ysum[k][d];
#pragma unroll
            for (int k = 0; k < U; k++) {
                if (idx[k] < N) mx = fmax(mx, ysum[k][d]);
                if (idx[k] < N && (long)idx[k] * D + d != skip) mn = fmin(mn, ysum[k][d]);
            }
            v[d] = wave_sum(t);
            v[D + 2 * d] = wave_max(mx);
            v[D + 2 * d + 1] = wave_min(mn);
        }
        const int w = threadIdx.x / kWave;
        __syncthreads();
        if (threadIdx.x % kWave == 0)
#pragma unroll
            for (int k = 0; k < 3 * D; k++) sh[k * NW + w] = v[k];
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                double r = 0.0, mx = sh[(D + 2 * d) * NW], mn = sh[(D + 2 * d + 1) * NW];
                for (int k = 0; k < NW; k++) r += sh[d * NW + k];
                for (int k = 1; k < NW; k++) {
                    mx = fmax(mx, sh[(D + 2 * d) * NW + k]);
                    mn = fmin(mn, sh[(D + 2 * d + 1) * NW + k]);
                }
                part[bid * D + d] = r;
                mmu[(bid * D + d) * 2 + 0] = mx;
                mmu[(bid * D + d) * 2 + 1] = mn;
            }
        }
    }
}

template <int P, int D>
__global__ __launch_bounds__(kThreads) void k_interp_f32(float *__restrict__ Y, const float *__restrict__ pos,
                                                         const float4 *__restrict__ gout,
                                                         const DevGeo *__restrict__ dg,
                                                         const double *__restrict__ zbuf,
                                                         const StepDev *__restrict__ sd,
                                                         float *__restrict__ dC, float *__restrict__ uY,
                                                         float *__restrict__ gains, double *__restrict__ part,
                                                         int N, double *__restrict__ mmu, long skip,
                                                         const double *__restrict__ ctr) {
    if (!dg->live) return;
    __shared__ double sh[3 * D * kThreads / kWave];
    interp_block<P, D>(Y, pos, gout, dg, zbuf, sd, dC, uY, gains, part, N, blockIdx.x, sh, mmu, skip, ctr);
}

// Mean of the updated Y from block partials (fixed order), FP64.  One workgroup of kThreads.
__device__ __forceinline__ void mean_block(const double *__restrict__ part, int nblk, int D, double N,
                                           double *__restrict__ mean, double *sh) {
    const int T = blockDim.x;
    // Each thread adds part[t], part[t + T], ... of a coordinate in that order.  The loads are issued
    // eight at a time; those past the end read as +0.0, which leaves the sum unchanged (it starts at
    // +0.0, so it is never -0.0).  In 2D both coordinates come from one 16-byte load.
    if (D == 2) {
        const double2 *p2 = reinterpret_cast<const double2 *>(part);
        double s0 = 0, s1 = 0;
        for (int b = threadIdx.x; b < nblk; b += 8 * T) {
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = b + u * T < nblk ? p2[b + u * T] : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                s0 += v[u].x;
                s1 += v[u].y;
            }
        }
        s0 = block_sum(s0, sh);
        if (threadIdx.x == 0) mean[0] = s0 / N;
        __syncthreads();
        s1 = block_sum(s1, sh);
        if (threadIdx.x == 0) mean[1] = s1 / N;
        return;
    }
    for (int d = 0; d < D; d++) {
        double s = 0;
        for (int b = threadIdx.x; b < nblk; b += 8 * T) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = b + u * T < nblk ? part[(long)(b + u * T) * D + d] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++) s += v[u];
        }
        s = block_sum(s, sh);
        if (threadIdx.x == 0) mean[d] = s / N;
        __syncthreads();
    }
}

__global__ void k_mean_final(const DevGeo *__restrict__ dg, const double *__restrict__ part, int nblk, int D,
                             double N, double *__restrict__ mean) {
    if (!dg->live) return;
    __shared__ double sh[kThreads / kWave];
    mean_block(part, nblk, D, N, mean, sh);
}

// k_mean_final, and the min / max that zeroMean will leave, for the host's next geometry: out = (max, min,
// Y[first]) of the centred Y, written to pinned host memory.  Centring maps each coordinate d through
// f_d(y) = (float)((double)y - mean_d), which never decreases, so the max and min of the centred values
// are f_d of the per-coordinate max and min that k_interp_f32 left in mmu (nblk blocks, D coordinates).
template <int D>
__global__ __launch_bounds__(kThreads) void k_mean_mm(const DevGeo *__restrict__ dg, const double *__restrict__ part,
                                                      const double *__restrict__ mmu, int nblk, double N,
                                                      const float *__restrict__ Y, long first,
                                                      double *__restrict__ mean, double *__restrict__ out) {
    if (!dg->live) return;
    __shared__ double sh[kThreads / kWave];
    mean_block(part, nblk, D, N, mean, sh);  // thread 0 wrote mean[0, D)
    __shared__ double shm[2 * kThreads / kWave];
    double gmx = -INFINITY, gmn = INFINITY;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double mx = -INFINITY, mn = INFINITY;
        const double2 *m2 = reinterpret_cast<const double2 *>(mmu) + d;
        for (int b = threadIdx.x; b < nblk; b += 4 * kThreads) {  // loads issued four at a time
            double2 u[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                u[k] = b + k * kThreads < nblk ? m2[(long)(b + k * kThreads) * D] : make_double2(-INFINITY, INFINITY);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                mx = fmax(mx, u[k].x);
                mn = fmin(mn, u[k].y);
            }
        }
        block_maxmin(mx, mn, shm);
        if (threadIdx.x == 0) {
            const double m = mean[d];
            gmx = fmax(gmx, (double)(float)(mx - m));
            gmn = fmin(gmn, (double)(float)(mn - m));
        }
    }
    if (threadIdx.x == 0) {
        out[0] = gmx;
        out[1] = gmn;
        out[2] = (double)(float)((double)Y[first] - mean[0]);
        __threadfence_system();
    }
}

// zeroMean (tsne.cpp:1851-1876): Y -= mean, with the min/max partials of the result for the next
// geometry (same layout as k_minmax_partial).
// Workgroup bid of nblk: the body of k_center and of its batched form.
template <typename TV>
__device__ __forceinline__ void center_block(TV *__restrict__ Y, long n, int D, long skip,
                                             const double *__restrict__ mean, double *__restrict__ part, int bid,
                                             int nblk) {
    double mx = -INFINITY, mn = INFINITY;
    const double m0 = mean[0], m1 = D == 2 ? mean[1] : 0.0;
    for (long e = (long)bid * kThreads + threadIdx.x; e < n; e += (long)nblk * kThreads) {
        const double m = D > 2 ? mean[e % D] : (D == 2 && (e & 1)) ? m1 : m0;
        const TV v = (TV)((double)Y[e] - m);
        Y[e] = v;
        mx = fmax(mx, (double)v);
        if (e != skip) mn = fmin(mn, (double)v);
    }
    __shared__ double sh[2][kThreads / kWave];
    mx = wave_max(mx);
    mn = wave_min(mn);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        sh[0][w] = mx;
        sh[1][w] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kThreads / kWave; i++) {
            mx = fmax(mx, sh[0][i]);
            mn = fmin(mn, sh[1][i]);
        }
        part[bid * 2 + 0] = mx;
        part[bid * 2 + 1] = mn;
    }
}

template <typename TV>
__global__ __launch_bounds__(kThreads) void k_center(const DevGeo *__restrict__ dg, TV *__restrict__ Y, long n,
                                                     int D, long skip, const double *__restrict__ mean,
                                                     double *__restrict__ part) {
    if (!dg->live) return;
    center_block<TV>(Y, n, D, skip, mean, part, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- spatial reordering (FP32 form)
// New row r is old row rel[r]; column indices are renumbered through inv (old -> new).  One
// wavefront per new row, entries copied in their original order (per-row sums keep their order).
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_permute_csr(const int64_t *__restrict__ rowP_old,
                                                          const uint32_t *__restrict__ colP_old,
                                                          const TV *__restrict__ val_old,
                                                          const int64_t *__restrict__ rowP_new,
                                                          const int *__restrict__ rel, const int *__restrict__ inv,
                                                          int N, uint32_t *__restrict__ colP_new,
                                                          TV *__restrict__ val_new) {
    const long r = ((long)blockIdx.x * kThreads + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (r >= N) return;
    const long o = rel[r];
    const long ob = rowP_old[o], ne = rowP_new[r + 1], nb = rowP_new[r];
    for (long k = lane; k < ne - nb; k += kWave) {
        colP_new[nb + k] = (uint32_t)inv[colP_old[ob + k]];
        val_new[nb + k] = val_old[ob + k];
    }
}

// out[r*D + d] = in[rel[r]*D + d]
template <typename TV>
__global__ __launch_bounds__(kThreads) void k_gather_rows(const TV *__restrict__ in, const int *__restrict__ rel, int N,
                                                          int D, TV *__restrict__ out) {
    const long e = (long)blockIdx.x * kThreads + threadIdx.x;
    if (e < (long)N * D) out[e] = in[(long)rel[e / D] * D + e % D];
}

// ---------------------------------------------------------------- plain FP64 path stages

template <typename T>
__global__ void k_fill(T *__restrict__ p, long n, T v) {
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads) p[e] = v;
}

// Charges of the reference (tsne.cpp:1057-1062, :904-908, :939-941, :786-790, :674-677, :689-691).
// set: 0 = [1, x, y, x^2+y^2] (2D df=1), 1 = [x, y, 1] (2D df), 2 = [1] (general kernel),
//      3 = [1, y, y^2] (1D df=1), 4 = [y, 1] (1D df).
template <int D>
__global__ __launch_bounds__(kThreads) void k_charges(const double *__restrict__ Y, int N, int set,
                                                      double *__restrict__ ch) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const double x = Y[(long)i * D], y = D == 2 ? Y[(long)i * D + 1] : 0.0;
    switch (set) {
    case 0: ch[4L * i] = 1; ch[4L * i + 1] = x; ch[4L * i + 2] = y; ch[4L * i + 3] = x * x + y * y; break;
    case 1: ch[3L * i] = x; ch[3L * i + 1] = y; ch[3L * i + 2] = 1; break;
    case 2: ch[i] = 1; break;
    case 3: ch[3L * i] = 1; ch[3L * i + 1] = x; ch[3L * i + 2] = x * x; break;
    default: ch[2L * i] = x; ch[2L * i + 1] = 1; break;
    }
}

// Box index, in-box coordinates and Lagrange weights per point (nbodyfft.cpp:85-127 / :349-371):
// box[i], wx[j*N + i], wy[j*N + i] (SoA like the reference's interpolated_values).
template <int P, int D>
__global__ __launch_bounds__(kThreads) void k_weights_f64(const double *__restrict__ xs, const double *__restrict__ ys,
                                                          int N, Geo g, const double *__restrict__ lox,
                                                          const double *__restrict__ loy, int *__restrict__ box,
                                                          double *__restrict__ wx, double *__restrict__ wy) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    double tx, ty = 0;
    int bx = box_coord<D>(xs[i], g, tx), by = 0;
    if (lox) tx = (xs[i] - lox[bx]) / g.bw;  // caller-given box bounds (mirror tier)
    if constexpr (D == 2) {
        by = box_coord<D>(ys[i], g, ty);
        if (loy) ty = (ys[i] - loy[by]) / g.bw;
    }
    box[i] = by * g.nb + bx;
    double L[P];
    lagrange<P, double>(tx, g, L);
#pragma unroll
    for (int j = 0; j < P; j++) wx[(long)j * N + i] = L[j];
    if constexpr (D == 2) {
        lagrange<P, double>(ty, g, L);
#pragma unroll
        for (int j = 0; j < P; j++) wy[(long)j * N + i] = L[j];
    }
}

// interpolate() of nbodyfft.cpp:310-336 on the GPU: out[j*N + i] = L_j(y[i]), spacings/denominators in g.
template <int P>
__global__ __launch_bounds__(kThreads) void k_lagrange_f64(const double *__restrict__ y, int N, Geo g,
                                                           double *__restrict__ out) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    double L[P];
    lagrange<P, double>(y[i], g, L);
#pragma unroll
    for (int j = 0; j < P; j++) out[(long)j * N + i] = L[j];
}

// Spread with FP64 atomics into w[node][nterms] (nbodyfft.cpp:129-147 / :373-382).
template <int P, int D>
__global__ __launch_bounds__(kThreads) void k_spread_f64(const int *__restrict__ box, const double *__restrict__ wx,
                                                         const double *__restrict__ wy,
                                                         const double *__restrict__ ch, int nterms, int N, Geo g,
                                                         double *__restrict__ w) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const int b = box[i];
    const int bx = b % g.nb, by = b / g.nb;
    for (int a = 0; a < P; a++) {
        const double lx = wx[(long)a * N + i];
        if constexpr (D == 2) {
            for (int c = 0; c < P; c++) {
                const double ly = wy[(long)c * N + i];
                const long idx = (long)(bx * P + a) * g.n1d + by * P + c;
                for (int d = 0; d < nterms; d++) atomicAdd(&w[idx * nterms + d], ly * lx * ch[(long)i * nterms + d]);
            }
        } else {
            const long idx = (long)bx * P + a;
            for (int d = 0; d < nterms; d++) atomicAdd(&w[idx * nterms + d], lx * ch[(long)i * nterms + d]);
        }
    }
}

// Interpolate potentials back (nbodyfft.cpp:222-239 / :439-447), accumulating like the reference.
template <int P, int D>
__global__ __launch_bounds__(kThreads) void k_interp_f64(const int *__restrict__ box, const double *__restrict__ wx,
                                                         const double *__restrict__ wy,
                                                         const double *__restrict__ V, int nterms, int N, Geo g,
                                                         double *__restrict__ pot) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const int b = box[i];
    const int bx = b % g.nb, by = b / g.nb;
    double acc[4] = {0, 0, 0, 0};
    for (int a = 0; a < P; a++) {
        const double lx = wx[(long)a * N + i];
        if constexpr (D == 2) {
            for (int c = 0; c < P; c++) {
                const double ly = wy[(long)c * N + i];
                const long idx = (long)(bx * P + a) * g.n1d + by * P + c;
                for (int d = 0; d < nterms; d++) acc[d] += lx * ly * V[idx * nterms + d];
            }
        } else {
            const long idx = (long)bx * P + a;
            for (int d = 0; d < nterms; d++) acc[d] += lx * V[idx * nterms + d];
        }
    }
    for (int d = 0; d < nterms; d++) pot[(long)i * nterms + d] += acc[d];
}

// Per-point Z terms and block partials (tsne.cpp:1101-1110 / :945-950 / :809-817 / :705-713).
// zform: 0 = 2D expansion on [1,x,y,x^2+y^2]; 1 = sum of phi (general kernel); 2 = 1D expansion.
__global__ __launch_bounds__(kThreads) void k_zterms_f64(const double *__restrict__ Y, const double *__restrict__ pot,
                                                         int N, int D, int zform, double *__restrict__ part) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    double s = 0;
    if (i < N) {
        if (zform == 0) {
            const double x = Y[2L * i], y = Y[2L * i + 1];
            const double *p = pot + 4L * i;
            s = (1 + x * x + y * y) * p[0] - 2 * (x * p[1] + y * p[2]) + p[3];
        } else if (zform == 1) {
            s = pot[i];
        } else {
            const double y = Y[i];
            const double *p = pot + 3L * i;
            s = (1 + y * y) * p[0] - 2 * (y * p[1]) + p[2];
        }
    }
    __shared__ double sh[kThreads / kWave];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum_final(const double *__restrict__ part, int nblk, double sub, double *__restrict__ out) {
    __shared__ double sh[kThreads / kWave];
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        out[0] = s - sub;
        out[1] = 1.0 / (s - sub);
    }
}

// dC = pos_f - neg_f (tsne.cpp:1148-1155 / :990-997 / :846-850 / :735-741).
// cform: 0 = 2D [1,x,y,..]: neg = (x phi0 - phi1, y phi0 - phi2) / Z, stride 4
//        1 = 2D df [x,y,1]: neg = (x phi2 - phi0, y phi2 - phi1) / Z, stride 3
//        2 = 1D [1,y,y^2]: neg = (y phi0 - phi1) / Z, stride 3;  3 = 1D df [y,1]: (y phi1 - phi0)/Z, stride 2
__global__ __launch_bounds__(kThreads) void k_combine_f64(const double *__restrict__ Y, const double *__restrict__ pos,
                                                          const double *__restrict__ pot, const double *__restrict__ zbuf,
                                                          int N, int cform, double *__restrict__ dC) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const double Z = zbuf[0];
    switch (cform) {
    case 0: {
        const double x = Y[2L * i], y = Y[2L * i + 1];
        const double *p = pot + 4L * i;
        dC[2L * i] = pos[2L * i] - (x * p[0] - p[1]) / Z;
        dC[2L * i + 1] = pos[2L * i + 1] - (y * p[0] - p[2]) / Z;
        break;
    }
    case 1: {
        const double x = Y[2L * i], y = Y[2L * i + 1];
        const double *p = pot + 3L * i;
        dC[2L * i] = pos[2L * i] - (x * p[2] - p[0]) / Z;
        dC[2L * i + 1] = pos[2L * i + 1] - (y * p[2] - p[1]) / Z;
        break;
    }
    case 2: {
        const double y = Y[i];
        dC[i] = pos[i] - (y * pot[3L * i] - pot[3L * i + 1]) / Z;
        break;
    }
    default: {
        const double y = Y[i];
        dC[i] = pos[i] - (y * pot[2L * i + 1] - pot[2L * i]) / Z;
        break;
    }
    }
}

// Optimiser update, FP64, exactly the reference's elementwise loops (tsne.cpp:478-514), with the
// block partial sums of the new Y for zeroMean.
__global__ __launch_bounds__(kThreads) void k_update_f64(double *__restrict__ Y, double *__restrict__ uY,
                                                         double *__restrict__ gains, const double *__restrict__ dY,
                                                         int N, int D, double momentum, double lr, double msn,
                                                         int plain_descent, double *__restrict__ part) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    double ys[4] = {0, 0, 0, 0};  // D <= 4
    if (i < N) {
        if (plain_descent) {
            for (int d = 0; d < D; d++) {
                Y[(long)i * D + d] = Y[(long)i * D + d] - dY[(long)i * D + d];
                ys[d] = Y[(long)i * D + d];
            }
        } else {
            double u[4], step = 0;
            for (int d = 0; d < D; d++) {
                const long e = (long)i * D + d;
                const double sd = dY[e] == .0 ? .0 : (dY[e] < .0 ? -1.0 : 1.0);
                const double su = uY[e] == .0 ? .0 : (uY[e] < .0 ? -1.0 : 1.0);
                double gn = sd != su ? gains[e] + .2 : gains[e] * .8;
                if (gn < .01) gn = .01;
                gains[e] = gn;
                u[d] = momentum * uY[e] - lr * gn * dY[e];
                step += u[d] * u[d];
            }
            if (msn > 0) {
                step = sqrt(step);
                if (step > msn)
                    for (int d = 0; d < D; d++) u[d] *= (msn / step);
            }
            for (int d = 0; d < D; d++) {
                const long e = (long)i * D + d;
                uY[e] = u[d];
                Y[e] = Y[e] + u[d];
                ys[d] = Y[e];
            }
        }
    }
    __shared__ double sh[kThreads / kWave];
    for (int d = 0; d < D; d++) {
        const double s = block_sum(ys[d], sh);
        if (threadIdx.x == 0) part[blockIdx.x * D + d] = s;
    }
}

// Split interleaved Y (FP64) into xs, ys (tsne.cpp:1042-1044).
__global__ void k_split_f64(const double *__restrict__ Y, int N, double *__restrict__ xs, double *__restrict__ ys) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < N) {
        xs[i] = Y[2L * i];
        ys[i] = Y[2L * i + 1];
    }
}

// ---------------------------------------------------------------- KL divergence over CSR edges
// C = sum_ij p_ij log((p_ij + FLT_MIN) / (q_ij / Z + FLT_MIN)), q = (1 / (1 + d^2 / df))^df
// (tsne.cpp:1329-1355), in FP64 with a fixed reduction order (no race).
template <typename TV, typename TI, typename TC, int D, int SW>
__global__ __launch_bounds__(kThreads) void k_kl(const TI *__restrict__ rowP, const TC *__restrict__ colP,
                                                 const TV *__restrict__ valP, double pscale, const TV *__restrict__ Y,
                                                 int N, double df, const double *__restrict__ zbuf,
                                                 double *__restrict__ part) {
    const long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    const long row = gid / SW;
    const int lane = (int)(gid % SW);
    double acc = 0;
    if (row < N) {
        const double Z = zbuf[0];
        TV yi[D];
        load_pt<TV, D>(Y, row, yi);
        for (long k = (long)rowP[row] + lane; k < (long)rowP[row + 1]; k += SW) {
            const long j = (long)colP[k];
            const double p = (double)valP[k] * pscale;
            TV yj[D];
            load_pt<TV, D>(Y, j, yj);
            double Q = 0;
            for (int d = 0; d < D; d++) {
                const double df_ = (double)yi[d] - (double)yj[d];
                Q += df_ * df_;
            }
            Q = (df == 1.0 ? 1.0 / (1.0 + Q) : pow(1.0 / (1.0 + Q / df), df)) / Z;
            acc += p * log((p + (double)FLT_MIN) / (Q + (double)FLT_MIN));
        }
    }
    __shared__ double sh[kThreads / kWave];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// ---------------------------------------------------------------- perplexity calibration (FP64)
// distances2similarities (tsne.cpp:1394-1469) for many points: one wavefront per point, lanes over
// the K distances; the bisection steps are the reference's (beta doubling/halving until bracketed,
// tol 1e-5, at most 200 evaluations).  Sums use a fixed wave reduction order.
// self_col != 0: row pt holds all N distances of point pt (K = N) and its own entry is DBL_MIN instead
// of exp(0), as the reference's `n >= 0` argument (tsne.cpp:1421, :1459).  accumulate != 0 adds the
// row to P instead of storing it (the perplexity-list sum, tsne.cpp:1489-1493).
__global__ __launch_bounds__(kThreads) void k_d2s(const double *__restrict__ Dist, double *__restrict__ P, int n,
                                                  int K, double perplexity, double sigma, int ifSquared,
                                                  const double *__restrict__ beta0, double *__restrict__ beta_out,
                                                  int *__restrict__ iters_out, int self_col, int accumulate) {
    const long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    const long pt = gid / kWave;
    const int lane = threadIdx.x % kWave;
    if (pt >= n) return;
    const double *Dp = Dist + pt * K;
    double *Pp = P + pt * K;
    double beta, beta_eval, sum_P;
    int iter = 0;
    if (perplexity > 0) {
        double min_beta = -DBL_MAX, max_beta = DBL_MAX;
        const double tol = 1e-5, logp = log(perplexity);
        beta = beta0 ? beta0[pt] : 1.0;
        beta_eval = beta;
        sum_P = DBL_MIN;
        bool found = false;
        while (!found && iter < 200) {
            double s = 0, hs = 0;
            for (int m = lane; m < K; m += kWave) {
                const double d = ifSquared ? Dp[m] : Dp[m] * Dp[m];
                const double pv = (self_col && m == pt) ? DBL_MIN : exp(-beta * d);
                s += pv;
                hs += beta * (d * pv);
            }
            s = wave_sum(s);
            hs = wave_sum(hs);
            beta_eval = beta;
            sum_P = DBL_MIN + s;
            const double H = (hs / sum_P) + log(sum_P);
            const double Hdiff = H - logp;
            if (Hdiff < tol && -Hdiff < tol) {
                found = true;
            } else if (Hdiff > 0) {
                min_beta = beta;
                beta = (max_beta == DBL_MAX || max_beta == -DBL_MAX) ? beta * 2.0 : (beta + max_beta) / 2.0;
            } else {
                max_beta = beta;
                beta = (min_beta == -DBL_MAX || min_beta == DBL_MAX) ? beta / 2.0 : (beta + min_beta) / 2.0;
            }
            iter++;
        }
    } else {
        beta = 1 / (2 * sigma * sigma);
        beta_eval = beta;
        double s = 0;
        for (int m = lane; m < K; m += kWave) {
            const double d = ifSquared ? Dp[m] : Dp[m] * Dp[m];
            s += (self_col && m == pt) ? DBL_MIN : exp(-beta * d);
        }
        sum_P = DBL_MIN + wave_sum(s);
    }
    // The reference normalises the P of the last evaluated beta (tsne.cpp:1466); when the loop ends
    // at its iteration cap, the returned beta has already moved past it.
    for (int m = lane; m < K; m += kWave) {
        const double d = ifSquared ? Dp[m] : Dp[m] * Dp[m];
        const double v = ((self_col && m == pt) ? DBL_MIN : exp(-beta_eval * d)) / sum_P;
        Pp[m] = accumulate ? Pp[m] + v : v;
    }
    if (lane == 0) {
        beta_out[pt] = beta;
        if (iters_out) iters_out[pt] = iter;
    }
}

}  // namespace fitsne
