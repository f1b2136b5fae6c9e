// fitsne_fft.hpp -- LDS-resident Stockham FFT for one workgroup, gfx950.
//
// A workgroup transforms `lines` complex lines of length n held in LDS (line l at buf + l*stride).
// n = 2^a 3^b is split into stages of radix 16, 8, 4, 2, 9 or 3; each radix-R butterfly is done in
// registers (R = 8, 16 by even/odd splitting down to radix 2; R = 9 as 3 x 3), so a length-1152
// line takes three stages.  Each stage reads all its inputs into registers, synchronises, then writes
// its outputs in place (Stockham autosort, natural order in and out):
//   in [j + r n/R],  twiddle W_{Ns R}^{(j mod Ns) r},  DFT_R,  out [(j - j mod Ns) R + j mod Ns + r Ns]
// Twiddles come from a table tw[e] = exp(-2 pi i e / n) built in FP64 on the device.
// Unnormalised; `inv` selects exp(+2 pi i ...).
#pragma once

#include "fitsne_common.hpp"

namespace fitsne {

// exp(-2 pi i k / R) for the in-register butterflies (k < R / 2 for R = 8, 16; k < 9 for R = 9)
__device__ __forceinline__ double twc(int R, int k) {
    // cos(2 pi k / R)
    constexpr double c16[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613};
    constexpr double c9[9] = {1.0, 0.76604444311897803520, 0.17364817766693034885, -0.5, -0.93969262078590838405,
                              -0.93969262078590838405, -0.5, 0.17364817766693034885, 0.76604444311897803520};
    return R == 9 ? c9[k] : c16[k * (16 / R)];
}
__device__ __forceinline__ double tws(int R, int k) {
    // sin(2 pi k / R)
    constexpr double s16[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                               1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173};
    constexpr double s9[9] = {0.0, 0.64278760968653932632, 0.98480775301220805936, 0.86602540378443864676,
                              0.34202014332566873304, -0.34202014332566873304, -0.86602540378443864676,
                              -0.98480775301220805936, -0.64278760968653932632};
    return R == 9 ? s9[k] : s16[k * (16 / R)];
}

// multiply by W_R^k (forward: exp(-2 pi i k/R); inverse: conjugate)
template <int R, int K, typename T, bool INV>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
    if constexpr (K == 0) {
        return a;
    } else if constexpr (R == 4 * K) {  // W = -i (forward)
        return INV ? mul_i(a) : mul_negi(a);
    } else {
        const T c = (T)twc(R, K), s = (T)tws(R, K);
        const T si = INV ? s : -s;
        return cx<T>{a.x * c - a.y * si, a.x * si + a.y * c};
    }
}

template <int R, typename T, bool INV>
__device__ __forceinline__ void dft_small(cx<T> (&v)[R]);

template <int H, int K, typename T, bool INV>
__device__ __forceinline__ void combine_step(cx<T> (&v)[2 * H], const cx<T> (&e)[H], const cx<T> (&o)[H]) {
    if constexpr (K < H) {
        const cx<T> t = rot<2 * H, K, T, INV>(o[K]);
        v[K] = e[K] + t;
        v[K + H] = e[K] - t;
        combine_step<H, K + 1, T, INV>(v, e, o);
    }
}

template <int R, typename T, bool INV>
__device__ __forceinline__ void dft_small(cx<T> (&v)[R]) {
    if constexpr (R == 2) {
        cx<T> a = v[0] + v[1], b = v[0] - v[1];
        v[0] = a;
        v[1] = b;
    } else if constexpr (R == 4) {
        cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        cx<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
        cx<T> jb1 = INV ? mul_i(b1) : mul_negi(b1);
        v[0] = a0 + b0;
        v[1] = a1 + jb1;
        v[2] = a0 - b0;
        v[3] = a1 - jb1;
    } else if constexpr (R == 3) {
        const T s3 = (T)0.86602540378443864676372317075294;
        cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
        cx<T> m = v[0] - (T)0.5 * t1;
        cx<T> s = s3 * t2;
        cx<T> rt = INV ? mul_i(s) : mul_negi(s);
        v[0] = v[0] + t1;
        v[1] = m + rt;
        v[2] = m - rt;
    } else if constexpr (R == 8 || R == 16) {
        constexpr int H = R / 2;
        cx<T> e[H], o[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        dft_small<H, T, INV>(e);
        dft_small<H, T, INV>(o);
        combine_step<H, 0, T, INV>(v, e, o);
    } else {  // R == 9: three DFT3 over m (index 3m + r), twiddle W9^{r k}, three DFT3 over r
        cx<T> a[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
#pragma unroll
            for (int m = 0; m < 3; m++) a[r][m] = v[3 * m + r];
            dft_small<3, T, INV>(a[r]);
        }
        // a[r][k] now holds sum_m v[3m+r] W3^{mk}; twiddle by W9^{r k}
        a[1][1] = rot<9, 1, T, INV>(a[1][1]);
        a[1][2] = rot<9, 2, T, INV>(a[1][2]);
        a[2][1] = rot<9, 2, T, INV>(a[2][1]);
        a[2][2] = rot<9, 4, T, INV>(a[2][2]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            cx<T> b[3] = {a[0][k], a[1][k], a[2][k]};
            dft_small<3, T, INV>(b);
#pragma unroll
            for (int s = 0; s < 3; s++) v[k + 3 * s] = b[s];
        }
    }
}

// LDS position of element e of the lines.  With PAD one slot is skipped after every 16 elements: the
// first stage writes R consecutive elements per lane with lanes R elements apart, which without the
// skip puts every lane of a write in the same bank (R = 16); with it they spread over the banks.  All
// other accesses are consecutive across lanes either way.
template <bool PAD>
__host__ __device__ __forceinline__ int lp(int e) { return PAD ? e + (e >> 4) : e; }
// LDS elements taken by n elements of lines
template <bool PAD>
__host__ __device__ __forceinline__ int lspan(int n) { return PAD ? n + (n >> 4) + 1 : n; }

// Quotient q / d for 0 <= q, d < 2^16 as one high multiply: with m = ceil(2^32 / d) the error of
// q m / 2^32 against q / d is q (m d - 2^32) / (d 2^32) < 1 / d, too small to cross the next integer.
struct DivU16 {
    unsigned m;
    int d;
};
__device__ __forceinline__ DivU16 divu16(int d) { return {d > 1 ? 0xffffffffu / (unsigned)d + 1u : 0u, d}; }
__device__ __forceinline__ int quot(int q, DivU16 v) { return v.d > 1 ? (int)__umulhi((unsigned)q, v.m) : q; }

template <typename T, int R, int NTHR, int MAXPER, bool PAD, bool INV>
__device__ __forceinline__ void fft_stage(cx<T> *buf, int n, int lines, int stride, int Ns,
                                          const cx<T> *__restrict__ tw) {
    const int nbf = n / R;
    const int total = nbf * lines;
    const int tstep = n / (Ns * R);
    const DivU16 dn = divu16(nbf), ds = divu16(Ns);
    cx<T> v[MAXPER][R];
#pragma unroll
    for (int c = 0; c < MAXPER; c++) {
        const int q = threadIdx.x + c * NTHR;
        if (q < total) {
            const int l = quot(q, dn), j = q - l * nbf;
            const int src = l * stride + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[c][r] = buf[lp<PAD>(src + r * nbf)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXPER; c++) {
        const int q = threadIdx.x + c * NTHR;
        if (q < total) {
            const int l = quot(q, dn), j = q - l * nbf;
            const int jm = j - quot(j, ds) * Ns;
            if (jm != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cx<T> w = tw[jm * r * tstep];
                    if (INV) w.y = -w.y;
                    v[c][r] = v[c][r] * w;
                }
            }
            dft_small<R, T, INV>(v[c]);
            const int dst = l * stride + (j - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) buf[lp<PAD>(dst + r * Ns)] = v[c][r];
        }
    }
    __syncthreads();
}

template <typename T, int NTHR, int MAXTOT, bool PAD, bool INV>
__device__ __forceinline__ void fft_lines_dir(cx<T> *buf, const FftPlan &plan, int lines, int stride,
                                              const cx<T> *__restrict__ tw) {
    static_assert(MAXTOT < 65536, "the stage index arithmetic (quot) needs indices below 2^16");
    constexpr int P16 = (MAXTOT / 16 + NTHR - 1) / NTHR;
    constexpr int P8 = (MAXTOT / 8 + NTHR - 1) / NTHR;
    constexpr int P4 = (MAXTOT / 4 + NTHR - 1) / NTHR;
    constexpr int P2 = (MAXTOT / 2 + NTHR - 1) / NTHR;
    constexpr int P9 = (MAXTOT / 9 + NTHR - 1) / NTHR;
    constexpr int P3 = (MAXTOT / 3 + NTHR - 1) / NTHR;
    int Ns = 1;
    for (int s = 0; s < plan.nst; s++) {
        const int R = plan.rad[s];
        switch (R) {
        case 16: fft_stage<T, 16, NTHR, P16, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        case 8: fft_stage<T, 8, NTHR, P8, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        case 4: fft_stage<T, 4, NTHR, P4, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        case 2: fft_stage<T, 2, NTHR, P2, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        case 9: fft_stage<T, 9, NTHR, P9, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        default: fft_stage<T, 3, NTHR, P3, PAD, INV>(buf, plan.n, lines, stride, Ns, tw); break;
        }
        Ns *= R;
    }
}

// Transform `lines` lines of length plan.n in place.  Requires plan.n * lines <= MAXTOT.  Inlined, so
// that the direction is a constant at every call and the compiler sees which address space buf and tw
// are in (LDS or global).
template <typename T, int NTHR, int MAXTOT, bool PAD = false>
__device__ __forceinline__ void fft_lines(cx<T> *buf, const FftPlan &plan, int lines, int stride,
                                          const cx<T> *__restrict__ tw, bool inv) {
    if (inv) fft_lines_dir<T, NTHR, MAXTOT, PAD, true>(buf, plan, lines, stride, tw);
    else fft_lines_dir<T, NTHR, MAXTOT, PAD, false>(buf, plan, lines, stride, tw);
}

}  // namespace fitsne
