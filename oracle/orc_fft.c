/*
 * orc_fft.c -- the oracle's own FFT, standing in for the five FFTW entry points the reference path
 * uses (fftw_plan_dft_r2c_2d / c2r_2d / dft_1d + execute, nbodyfft.cpp:64-68, :164-168, :181, :194,
 * :301-304, :400-404, :410, :423).  TEST INFRASTRUCTURE: see fitsne_oracle.h.
 *
 * Mixed-radix decimation in time over the prime factors of n (radix 4 and 2 in closed form, any
 * other factor as a direct length-p DFT), unnormalised, sign -1 forward.  Pinned in tests against
 * orc_dft_c2c_naive (the DFT definition evaluated in long double).
 */
#include "fitsne_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

typedef struct {
    int n;
    int nfac;
    int fac[64];
    cplx *tw; /* tw[j] = exp(-2 pi i j / n) */
} fft_plan;

static void plan_init(fft_plan *pl, int n) {
    pl->n = n;
    pl->nfac = 0;
    int m = n;
    while (m % 4 == 0) { pl->fac[pl->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->fac[pl->nfac++] = 2; m /= 2; }
    for (int p = 3; (long)p * p <= m; p += 2)
        while (m % p == 0) { pl->fac[pl->nfac++] = p; m /= p; }
    if (m > 1) pl->fac[pl->nfac++] = m;
    if (pl->nfac == 0) pl->fac[pl->nfac++] = 1;
    pl->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; j++) {
        /* angle reduced to the first half-turn symmetric form for accuracy */
        double a = -2.0 * M_PI * (double)j / (double)n;
        pl->tw[j] = cos(a) + I * sin(a);
    }
}

static void plan_free(fft_plan *pl) { free(pl->tw); }

static inline cplx twid(const fft_plan *pl, long e, int inverse) {
    cplx w = pl->tw[e % pl->n];
    return inverse ? conj(w) : w;
}

/* out[0..n) = DFT of in[0], in[istride], ... ; factors fac[0..] multiply to n. */
static void fft_rec(cplx *out, const cplx *in, int n, long istride, const int *fac, const fft_plan *pl,
                    int inverse, cplx *scratch) {
    int p = fac[0];
    int m = n / p;
    if (m == 1) {
        for (int q = 0; q < p; q++) out[q] = in[q * istride];
    } else {
        for (int q = 0; q < p; q++) fft_rec(out + (long)q * m, in + q * istride, m, istride * p, fac + 1, pl, inverse, scratch);
    }
    if (p == 1) return;
    long tstep = pl->n / n; /* W_n^e  = W_N^(e*tstep) */
    long pstep = pl->n / p; /* W_p^e  = W_N^(e*pstep) */
    for (int k = 0; k < m; k++) {
        cplx *t = scratch;
        for (int q = 0; q < p; q++) {
            cplx v = out[(long)q * m + k];
            t[q] = (q == 0 || k == 0) ? v : v * twid(pl, (long)q * k * tstep, inverse);
        }
        if (p == 2) {
            out[k] = t[0] + t[1];
            out[m + k] = t[0] - t[1];
        } else if (p == 4) {
            /* W_4 = -i forward, +i inverse */
            cplx a0 = t[0] + t[2], a1 = t[0] - t[2];
            cplx b0 = t[1] + t[3], b1 = t[1] - t[3];
            cplx jb1 = inverse ? (I * b1) : (-I * b1);
            out[k] = a0 + b0;
            out[m + k] = a1 + jb1;
            out[2 * (long)m + k] = a0 - b0;
            out[3 * (long)m + k] = a1 - jb1;
        } else {
            for (int s = 0; s < p; s++) {
                cplx acc = t[0];
                for (int q = 1; q < p; q++) acc += t[q] * twid(pl, ((long)q * s % p) * pstep, inverse);
                out[(long)s * m + k] = acc;
            }
        }
    }
}

static int max_factor(const fft_plan *pl) {
    int mx = 1;
    for (int i = 0; i < pl->nfac; i++) if (pl->fac[i] > mx) mx = pl->fac[i];
    return mx;
}

static void fft_exec(const fft_plan *pl, const cplx *in, cplx *out, int inverse) {
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * (size_t)max_factor(pl));
    if (in == out) {
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)pl->n);
        memcpy(tmp, in, sizeof(cplx) * (size_t)pl->n);
        fft_rec(out, tmp, pl->n, 1, pl->fac, pl, inverse, scratch);
        free(tmp);
    } else {
        fft_rec(out, in, pl->n, 1, pl->fac, pl, inverse, scratch);
    }
    free(scratch);
}

void orc_fft_c2c(int n, const double *in, double *out, int inverse) {
    fft_plan pl;
    plan_init(&pl, n);
    fft_exec(&pl, (const cplx *)in, (cplx *)out, inverse);
    plan_free(&pl);
}

void orc_dft_c2c_naive(int n, const double *in, double *out, int inverse) {
    long double _Complex *acc = (long double _Complex *)malloc(sizeof(long double _Complex) * (size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; k++) {
        long double _Complex s = 0;
        for (int j = 0; j < n; j++) {
            long e = ((long)j * k) % n;
            long double a = (inverse ? 2.0L : -2.0L) * pi * (long double)e / (long double)n;
            long double _Complex x = (long double)in[2 * j] + I * (long double)in[2 * j + 1];
            s += x * (cosl(a) + I * sinl(a));
        }
        acc[k] = s;
    }
    for (int k = 0; k < n; k++) {
        out[2 * k] = (double)creall(acc[k]);
        out[2 * k + 1] = (double)cimagl(acc[k]);
    }
    free(acc);
}

/* Real n0 x n1 (row-major) -> complex n0 x (n1/2+1), as fftw_plan_dft_r2c_2d. */
void orc_fft_r2c_2d(int n0, int n1, const double *in, double *out_) {
    int h = n1 / 2 + 1;
    cplx *out = (cplx *)out_;
    fft_plan p1, p0;
    plan_init(&p1, n1);
    plan_init(&p0, n0);
    cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
    cplx *rowf = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
    for (int r = 0; r < n0; r++) {
        for (int j = 0; j < n1; j++) row[j] = in[(long)r * n1 + j];
        fft_exec(&p1, row, rowf, 0);
        for (int k = 0; k < h; k++) out[(long)r * h + k] = rowf[k];
    }
    free(row);
    free(rowf);
    cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)n0);
    cplx *colf = (cplx *)malloc(sizeof(cplx) * (size_t)n0);
    for (int k = 0; k < h; k++) {
        for (int r = 0; r < n0; r++) col[r] = out[(long)r * h + k];
        fft_exec(&p0, col, colf, 0);
        for (int r = 0; r < n0; r++) out[(long)r * h + k] = colf[r];
    }
    free(col);
    free(colf);
    plan_free(&p1);
    plan_free(&p0);
}

/* Complex n0 x (n1/2+1) Hermitian half -> real n0 x n1, unnormalised, as fftw_plan_dft_c2r_2d.
 * The input is not modified (FFTW may destroy it; the reference never reuses it). */
void orc_fft_c2r_2d(int n0, int n1, const double *in_, double *out) {
    int h = n1 / 2 + 1;
    const cplx *in = (const cplx *)in_;
    fft_plan p1, p0;
    plan_init(&p1, n1);
    plan_init(&p0, n0);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n0 * h);
    cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)n0);
    cplx *colf = (cplx *)malloc(sizeof(cplx) * (size_t)n0);
    for (int k = 0; k < h; k++) {
        for (int r = 0; r < n0; r++) col[r] = in[(long)r * h + k];
        fft_exec(&p0, col, colf, 1);
        for (int r = 0; r < n0; r++) tmp[(long)r * h + k] = colf[r];
    }
    free(col);
    free(colf);
    cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
    cplx *rowf = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
    for (int r = 0; r < n0; r++) {
        const cplx *src = tmp + (long)r * h;
        for (int k = 0; k < h; k++) row[k] = src[k];
        for (int k = h; k < n1; k++) row[k] = conj(src[n1 - k]);
        fft_exec(&p1, row, rowf, 1);
        for (int j = 0; j < n1; j++) out[(long)r * n1 + j] = creal(rowf[j]);
    }
    free(row);
    free(rowf);
    free(tmp);
    plan_free(&p1);
    plan_free(&p0);
}
