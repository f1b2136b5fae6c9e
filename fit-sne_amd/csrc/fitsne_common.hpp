// fitsne_common.hpp -- device types and helpers shared by the FIt-SNE gradient kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fitsne {

constexpr int kWave = 64;   // CDNA wavefront
constexpr int kMaxP = 5;    // largest n_interpolation_points with a compiled kernel

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// multiply by +i or -i
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return {-a.y, a.x}; }
template <typename T>
__device__ __forceinline__ cx<T> mul_negi(cx<T> a) { return {a.y, -a.x}; }

// Grid geometry of one gradient evaluation, from (min, max) exactly as tsne.cpp:1042-1092 /
// nbodyfft.cpp:15-46 (2D) and tsne.cpp:656-663 / nbodyfft.cpp:257-280 (1D).  Computed by the same
// functions on the host (one-shot calls) and on the device (captured steps).
struct Geo {
    double mn;       // coord_min = box_lower_bounds[0]
    double bw0;      // box width used for box bounds and nodes: (max - min) / n_boxes
    double bw;       // box width recomputed by n_body_fft: upper[0] - lower[0] (nbodyfft.cpp:79-80)
    double h;        // node spacing (1 / p) * bw0
    double sp[kMaxP];   // in-box node positions (nbodyfft.cpp:30-34)
    double den[kMaxP];  // Lagrange denominators (nbodyfft.cpp:313-321)
    float rdenf[kMaxP]; // 1 / den rounded once to FP32 (the FP32 form multiplies instead of dividing)
    int nb;          // boxes per dimension
    int n1d;         // nodes per dimension = p * nb
    int Mp;          // FFT length (2^a 3^b >= 2 n1d - 1)
    int H;           // Mp / 2 + 1
    int nbox;        // nb^D
    int D;
    int p;
    double fix;      // fixed-point scale of the FP32 spread's integer grid (spread_fix_scale)
    int fix_fast;    // every run sum of 64 points stays below 2^50 at that scale (spread_fix_fast)
};

// Stage radices of one FFT length (Stockham).
struct FftPlan {
    int n;
    int nst;
    int rad[24];
};

// Geometry and plan of the current step as the kernels read them (device memory).  `live` is 0 when
// the step's grid exceeds the build's limits: every kernel of the step then returns at once and the
// state is left as it was (the host reports the error).
struct DevGeo {
    Geo g;
    FftPlan pl;
    int live;
    int kspec_stale;  // kernel spectra must be recomputed this step
    int tw_stale;     // twiddles must be recomputed this step
    int err;          // set when a step was refused for its grid size
    double kc_mn, kc_mx;  // (min, max, nb) the kernel spectra were computed for
    int kc_nb;
    int tw_n;         // FFT length of the twiddle table
    double mn, mx;    // the step's (reference-rule) min and max
};

// Optimiser parameters of one captured step (tsne.cpp:478-514) and the P multiplier.
struct StepDev {
    float momentum, lr, max_step_norm;
    int plain_descent;
    int mode;       // 0: gradient only; 1: update
    float exag;
};

// tsne.cpp:1065-1077
__host__ __device__ inline int n_boxes_2d(double mn, double mx, double ipi, int mni) {
    int nb = (int)fmax((double)mni, (mx - mn) / ipi);
    const int allowed[20] = {25, 36, 50, 55, 60, 65, 70, 75, 80, 85, 90, 96, 100, 110, 120, 130, 140, 150, 175, 200};
    if (nb < allowed[19]) {
        int c = 0;
        while (allowed[c] < nb) c++;
        nb = allowed[c];
    }
    return nb;
}

// tsne.cpp:663
__host__ __device__ inline int n_boxes_1d(double mn, double mx, double ipi, int mni) {
    return (int)fmax((double)mni, (mx - mn) / ipi);
}

// Barrier-separated stages of make_plan's radices for length m = 2^a 3^b.
__host__ __device__ inline int fft_stages(long m) {
    int s = 0;
    while (m % 16 == 0) { s++; m /= 16; }
    if (m % 8 == 0) { s++; m /= 8; }
    if (m % 4 == 0) { s++; m /= 4; }
    if (m % 2 == 0) { s++; m /= 2; }
    while (m % 9 == 0) { s++; m /= 9; }
    while (m % 3 == 0) { s++; m /= 3; }
    return s;
}

// FFT length for n = 2 n1d - 1 (any length >= n gives the same linear convolution): the smallest
// 2^a 3^b >= n; for short lines (up to 512) the 2^a 3^b in [n, 2^ceil(log2 n)] with the least
// length x stages, since there a stage's barriers cost more than its points (486 = 2 x 9 x 9 x 3 became
// 512 = 16 x 16 x 2: the 64-problem batch 46,700 -> 48,650 problem-iterations/s; for longer lines,
// 972 -> 1024 and 3888 -> 4096, the step did not gain or lost 1%).  The choice never passes the next
// power of two, so a grid within a power-of-two length limit stays within it.
constexpr int kFftShortLen = 512;
__host__ __device__ inline int fft_size(int n) {
    long pw = 1;
    while (pw < n) pw *= 2;
    long best = -1, bc = 0;
    for (long p3 = 1; p3 <= pw; p3 *= 3) {
        long m = p3;
        while (m < n) m *= 2;
        if (m > pw) continue;
        const long cost = pw <= kFftShortLen ? m * fft_stages(m) : m;
        if (best < 0 || cost < bc || (cost == bc && m < best)) {
            best = m;
            bc = cost;
        }
    }
    return (int)best;
}

// Stage radices: 16, 8, 4, 2 for the power of two, then 9, 3 (fewest barrier-separated stages).
__host__ __device__ inline FftPlan make_plan(int n) {
    FftPlan pl{};
    pl.n = n;
    int m = n;
    while (m % 16 == 0) { pl.rad[pl.nst++] = 16; m /= 16; }
    if (m % 8 == 0) { pl.rad[pl.nst++] = 8; m /= 8; }
    if (m % 4 == 0) { pl.rad[pl.nst++] = 4; m /= 4; }
    if (m % 2 == 0) { pl.rad[pl.nst++] = 2; m /= 2; }
    while (m % 9 == 0) { pl.rad[pl.nst++] = 9; m /= 9; }
    while (m % 3 == 0) { pl.rad[pl.nst++] = 3; m /= 3; }
    return pl;  // m == 1 by construction of fft_size
}

// Node spacings and Lagrange denominators as nbodyfft.cpp:30-34 and :313-321.
__host__ __device__ inline void fill_spacings(Geo &g, int p) {
    double h = 1 / (double)p;
    g.sp[0] = h / 2;
    for (int i = 1; i < p; i++) g.sp[i] = g.sp[i - 1] + h;
    for (int i = 0; i < p; i++) {
        g.den[i] = 1;
        for (int j = 0; j < p; j++)
            if (i != j) g.den[i] *= g.sp[i] - g.sp[j];
        g.rdenf[i] = (float)(1.0 / g.den[i]);
    }
}

// Geometry of one gradient from the (reference-rule) min and max, as precompute_2d / precompute
// compute it (nbodyfft.cpp:15-46, :257-280) and n_body_fft re-derives the box width (:79-80, :344-345).
__host__ __device__ inline Geo make_geo(int D, int p, double mn, double mx, int nb) {
    Geo g{};
    g.D = D;
    g.p = p;
    g.nb = nb;
    g.mn = mn;
    g.bw0 = (mx - mn) / (double)nb;
    const double lower0 = 0 * g.bw0 + mn, upper0 = (0 + 1) * g.bw0 + mn;
    g.bw = upper0 - lower0;
    g.h = (1 / (double)p) * g.bw0;
    fill_spacings(g, p);
    g.n1d = p * nb;
    g.Mp = fft_size(2 * g.n1d - 1);
    g.H = g.Mp / 2 + 1;
    g.nbox = D == 2 ? nb * nb : nb;
    return g;
}

// Largest |L_j(t)| over the nodes j, at t (in box widths).
__host__ __device__ inline double lagrange_absmax(const Geo &g, double t) {
    double m = 0;
    for (int j = 0; j < g.p; j++) {
        double v = 1;
        for (int k = 0; k < g.p; k++)
            if (k != j) v *= t - g.sp[k];
        m = fmax(m, fabs(v / g.den[j]));
    }
    return m;
}

// Largest in-box Lagrange weight, max over j and t in [0, 1] of |L_j(t)| (sampled at 257 points, with a
// 10% margin that covers the variation of these low-degree polynomials between samples).  It depends
// only on p: computed once per context on the host and passed to the device geometry.
inline double lagrange_inbox_bound(int p) {
    Geo g{};
    g.p = p;
    fill_spacings(g, p);
    double lb = 0;
    for (int s = 0; s <= 256; s++) lb = fmax(lb, lagrange_absmax(g, s / 256.0));
    return 1.1 * lb;
}

// Fixed-point scale of the FP32 spread's int64 grid: each run of points' sums is rounded to an integer
// at this scale and the integers add exactly, so no node's total may reach 2^62.  A bound on that total:
// every point of the step in one node at the largest in-box Lagrange weight lb (lagrange_inbox_bound),
// times D for the |e|^2 moment, plus the one point the 2D min rule can leave outside the grid (the first
// x coordinate, `gap` box widths below box 0, with its extrapolated weights).  2^38 while that bound
// stays below 2^24; above, the largest power of two that keeps the total under 2^62.  0 (refuse the
// step) when that is below 2^20: the outlier's weights would swamp every other point's resolution.
// Evaluated without contraction, so that host and device agree bit for bit.
__host__ __device__ inline double spread_fix_scale(const Geo &g, double n_points, double gap, double lb) {
#pragma clang fp contract(off)
    double bound = n_points * (g.D == 2 ? lb * lb * 2 : lb);
    if (g.D == 2 && gap > 0) bound += 1.1 * lagrange_absmax(g, -gap) * lb * ((1 + gap) * (1 + gap) + 1);
    if (!(bound < 16777216.0)) {          // 2^24
        if (!(bound < 4.0e18)) return 0;  // also a NaN bound
        int e = 0;
        (void)frexp(bound, &e);  // bound < 2^e
        if (62 - e < 20) return 0;
        return ldexp(1.0, 62 - e);
    }
    return 274877906944.0;  // 2^38
}

// Whether the spread may round its run sums with the 2^51-range rounding alone: a wavefront's run holds at
// most 64 points, each with every value below lb^D D (in the box) or the extrapolated first point's bound,
// so the run sums stay below 2^50 at scale fix unless that outlier is far out.
__host__ __device__ inline int spread_fix_fast(const Geo &g, double gap, double lb) {
#pragma clang fp contract(off)
    double per = g.D == 2 ? lb * lb * 2 : lb;
    if (g.D == 2 && gap > 0) per = fmax(per, 1.1 * lagrange_absmax(g, -gap) * lb * ((1 + gap) * (1 + gap) + 1));
    return 64.0 * per * g.fix < 1125899906842624.0;  // 2^50
}

// Workgroup-to-row map over npad = 8 ceil(n / 8) slots: consecutive workgroups go to different XCDs
// (round-robin dispatch), so each XCD takes a contiguous range of rows and its L2 merges their
// lines.  A bijection of [0, npad); slots mapping to rows >= n are skipped by the caller.
__device__ __forceinline__ int xcd_slot_row(int slot, int n) {
    const int per = (n + 7) / 8;
    return (slot % 8) * per + slot / 8;
}

// ---- reductions ----
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
template <typename T, int W>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// Block-wide sum of one double for blockDim.x <= 1024 (result valid in thread 0).
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); i++) r += sh[i];
    return r;
}

// Block-wide max and min (results valid in thread 0); sh holds 2 blockDim.x / kWave doubles.
__device__ __forceinline__ void block_maxmin(double &mx, double &mn, double *sh) {
    mx = wave_max(mx);
    mn = wave_min(mn);
    const int nw = blockDim.x / kWave, w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        sh[w] = mx;
        sh[nw + w] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < nw; i++) {
            mx = fmax(mx, sh[i]);
            mn = fmin(mn, sh[nw + i]);
        }
}

// Lagrange basis on p equispaced in-box nodes (the reference's interpolate(), nbodyfft.cpp:322-333):
// L_j(t) = prod_{k != j} (t - s_k) / den_j.
template <int P, typename T>
__device__ __forceinline__ void lagrange(T t, const Geo &g, T (&L)[P]) {
#pragma unroll
    for (int j = 0; j < P; j++) {
        T v = (T)1;
#pragma unroll
        for (int k = 0; k < P; k++)
            if (k != j) v *= t - (T)g.sp[k];
        L[j] = v / (T)g.den[j];
    }
}

// Box index along one axis and the in-box coordinate, in FP64 exactly as nbodyfft.cpp:86-113 (2D,
// clamp both sides) or :350-363 (1D, clamp above only).
// The reference truncates toward zero and clamps; written on the quotient so that a NaN or an
// out-of-range value can never produce an index outside [0, nb) (identical for finite inputs: a
// quotient in (-1, 0) truncates to 0 either way, and the 1D path only sees quotients >= 0).
template <int D>
__device__ __forceinline__ int box_coord(double x, const Geo &g, double &t) {
    const double q = (x - g.mn) / g.bw;
    int b;
    if (!(q >= 0.0)) b = 0;
    else if (q >= (double)g.nb) b = g.nb - 1;
    else b = (int)q;
    t = (x - (b * g.bw0 + g.mn)) / g.bw;
    return b;
}

}  // namespace fitsne
