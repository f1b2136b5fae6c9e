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

// Grid geometry of one gradient evaluation, computed on the host from (min, max) exactly as
// tsne.cpp:1042-1092 / nbodyfft.cpp:15-46 (2D) and tsne.cpp:656-663 / nbodyfft.cpp:257-280 (1D).
struct Geo {
    double mn;       // coord_min = box_lower_bounds[0]
    double bw0;      // box width used for box bounds and nodes: (max - min) / n_boxes
    double bw;       // box width recomputed by n_body_fft: upper[0] - lower[0] (nbodyfft.cpp:79-80)
    double h;        // node spacing (1 / p) * bw0
    double sp[kMaxP];   // in-box node positions (nbodyfft.cpp:30-34)
    double den[kMaxP];  // Lagrange denominators (nbodyfft.cpp:313-321)
    int nb;          // boxes per dimension
    int n1d;         // nodes per dimension = p * nb
    int Mp;          // FFT length (2^a 3^b >= 2 n1d - 1)
    int H;           // Mp / 2 + 1
    int nbox;        // nb^D
    int D;
    int p;
};

// Stage radices of one FFT length (Stockham, radices 4/2/3).
struct FftPlan {
    int n;
    int nst;
    int rad[24];
};

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
