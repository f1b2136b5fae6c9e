// fitsne_fft.hpp -- LDS-resident Stockham FFT for one workgroup (radices 4, 2, 3), gfx950.
//
// A workgroup transforms `lines` complex lines of length n held in LDS (line l at buf + l*stride).
// Each stage of radix R reads all its inputs into registers, synchronises, then writes the
// butterflies' outputs in place (Stockham autosort, natural order in and out):
//   in  [j + r n/R],  twiddle W_{Ns R}^{(j mod Ns) r},  DFT_R,  out [(j - j mod Ns) R + j mod Ns + r Ns]
// Twiddles come from a table tw[e] = exp(-2 pi i e / n) built in FP64 on the device.
// Unnormalised; `inv` selects exp(+2 pi i ...).
#pragma once

#include "fitsne_common.hpp"

namespace fitsne {

template <int R, typename T>
__device__ __forceinline__ void dft_small(cx<T> (&v)[R], bool inv) {
    if constexpr (R == 2) {
        cx<T> a = v[0] + v[1], b = v[0] - v[1];
        v[0] = a;
        v[1] = b;
    } else if constexpr (R == 4) {
        cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        cx<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
        cx<T> jb1 = inv ? mul_i(b1) : mul_negi(b1);
        v[0] = a0 + b0;
        v[1] = a1 + jb1;
        v[2] = a0 - b0;
        v[3] = a1 - jb1;
    } else {  // R == 3
        const T s3 = (T)0.86602540378443864676372317075294;
        cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
        cx<T> m = v[0] - (T)0.5 * t1;
        cx<T> s = s3 * t2;
        cx<T> rot = inv ? mul_i(s) : mul_negi(s);
        v[0] = v[0] + t1;
        v[1] = m + rot;
        v[2] = m - rot;
    }
}

template <typename T, int R, int NTHR, int MAXPER>
__device__ __forceinline__ void fft_stage(cx<T> *buf, int n, int lines, int stride, int Ns,
                                          const cx<T> *__restrict__ tw, bool inv) {
    const int nbf = n / R;
    const int total = nbf * lines;
    const int tstep = n / (Ns * R);
    cx<T> v[MAXPER][R];
#pragma unroll
    for (int c = 0; c < MAXPER; c++) {
        const int q = threadIdx.x + c * NTHR;
        if (q < total) {
            const int l = q / nbf, j = q - l * nbf;
            const cx<T> *src = buf + l * stride + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[c][r] = src[r * nbf];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXPER; c++) {
        const int q = threadIdx.x + c * NTHR;
        if (q < total) {
            const int l = q / nbf, j = q - l * nbf;
            const int jm = j % Ns;
            if (jm != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cx<T> w = tw[jm * r * tstep];
                    if (inv) w.y = -w.y;
                    v[c][r] = v[c][r] * w;
                }
            }
            dft_small<R, T>(v[c], inv);
            cx<T> *dst = buf + l * stride + (j - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * Ns] = v[c][r];
        }
    }
    __syncthreads();
}

// Transform `lines` lines of length plan.n in place.  Requires plan.n * lines <= MAXTOT.
template <typename T, int NTHR, int MAXTOT>
__device__ void fft_lines(cx<T> *buf, const FftPlan &plan, int lines, int stride, const cx<T> *__restrict__ tw,
                          bool inv) {
    constexpr int P4 = (MAXTOT / 4 + NTHR - 1) / NTHR;
    constexpr int P2 = (MAXTOT / 2 + NTHR - 1) / NTHR;
    constexpr int P3 = (MAXTOT / 3 + NTHR - 1) / NTHR;
    int Ns = 1;
    for (int s = 0; s < plan.nst; s++) {
        const int R = plan.rad[s];
        if (R == 4) fft_stage<T, 4, NTHR, P4>(buf, plan.n, lines, stride, Ns, tw, inv);
        else if (R == 2) fft_stage<T, 2, NTHR, P2>(buf, plan.n, lines, stride, Ns, tw, inv);
        else fft_stage<T, 3, NTHR, P3>(buf, plan.n, lines, stride, Ns, tw, inv);
        Ns *= R;
    }
}

}  // namespace fitsne
