/*
 * orc_tsne.c -- restatement of the gradient orchestrators, the optimiser step, the FFT-path KL and
 * the affinity code of src/tsne.cpp (FIt-SNE 1.2.1).  TEST INFRASTRUCTURE: see fitsne_oracle.h.
 */
#include "fitsne_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- grid geometry ---- */

/* tsne.cpp:1042-1049: `else if` -- a value that sets a new running max is never tested against min. */
void orc_minmax_2d_quirk(const double *Y, int N, double *min_coord, double *max_coord) {
    double mn = INFINITY, mx = -INFINITY;
    for (long i = 0; i < N; i++) {
        double x = Y[i * 2 + 0], y = Y[i * 2 + 1];
        if (x > mx) mx = x;
        else if (x < mn) mn = x;
        if (y > mx) mx = y;
        else if (y < mn) mn = y;
    }
    *min_coord = mn;
    *max_coord = mx;
}

/* tsne.cpp:1065-1077 */
int orc_n_boxes_2d(double min_coord, double max_coord, double intervals_per_integer, int min_num_intervals) {
    int nb = (int)fmax(min_num_intervals, (max_coord - min_coord) / intervals_per_integer);
    static const int allowed[20] = {25, 36, 50, 55, 60, 65, 70, 75, 80, 85, 90, 96, 100, 110, 120, 130, 140, 150, 175, 200};
    if (nb < allowed[19]) {
        int c;
        for (c = 0; allowed[c] < nb; c++) {}
        nb = allowed[c];
    }
    return nb;
}

/* tsne.cpp:663 / :774 (no rounding table in 1D) */
int orc_n_boxes_1d(double y_min, double y_max, double intervals_per_integer, int min_num_intervals) {
    return (int)fmax(min_num_intervals, (y_max - y_min) / intervals_per_integer);
}

/* ---- attractive forces (PARALLEL_FOR over CSR rows) ---- */
typedef struct {
    const uint32_t *row_P, *col_P;
    const double *val_P, *Y;
    int D;
    double df;
    double *pos_f;
} att_ctx;

static void att_body(int lo, int hi, void *c_) {
    att_ctx *c = (att_ctx *)c_;
    const double *Y = c->Y;
    if (c->D == 2) {
        for (int n = lo; n < hi; n++) {
            double dim1 = 0, dim2 = 0;
            double xn = Y[2 * (long)n], yn = Y[2 * (long)n + 1];
            for (uint32_t i = c->row_P[n]; i < c->row_P[n + 1]; i++) {
                uint32_t m = c->col_P[i];
                double dx = xn - Y[2 * (long)m], dy = yn - Y[2 * (long)m + 1];
                double d_ij = dx * dx + dy * dy;
                double q_ij = c->df == 1.0 ? 1 / (1 + d_ij) : 1 / (1 + d_ij / c->df);
                dim1 += c->val_P[i] * q_ij * dx;
                dim2 += c->val_P[i] * q_ij * dy;
            }
            c->pos_f[2 * (long)n] = dim1;
            c->pos_f[2 * (long)n + 1] = dim2;
        }
    } else {
        for (int n = lo; n < hi; n++) {
            double dim1 = 0;
            for (uint32_t i = c->row_P[n]; i < c->row_P[n + 1]; i++) {
                double d_ij = Y[n] - Y[c->col_P[i]];
                double q_ij = c->df == 1.0 ? 1 / (1 + d_ij * d_ij) : 1 / (1 + (d_ij * d_ij) / c->df);
                dim1 += c->val_P[i] * q_ij * d_ij;
            }
            c->pos_f[n] = dim1;
        }
    }
}

/* tsne.cpp:1121-1137 (2D), :965-980 (2D df), :826-837 (1D), :721-732 (1D df) */
void orc_attractive(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y,
                    int N, int D, double df, unsigned nthreads, double *pos_f) {
    att_ctx c = {row_P, col_P, val_P, Y, D, df, pos_f};
    orc_parallel_for(nthreads, N, att_body, &c);
}

/* ---- tsne.cpp:1027-1169  computeFftGradient (2D, df = 1) ---- */
void orc_grad_2d(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                 double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double *sum_Q_out) {
    for (long i = 0; i < (long)N * 2; i++) dC[i] = 0.0;
    double *xs = (double *)malloc(sizeof(double) * (size_t)N);
    double *ys = (double *)malloc(sizeof(double) * (size_t)N);
    for (long i = 0; i < N; i++) { xs[i] = Y[i * 2]; ys[i] = Y[i * 2 + 1]; }
    double min_coord, max_coord;
    orc_minmax_2d_quirk(Y, N, &min_coord, &max_coord);

    int n_terms = 4;
    double *charges = (double *)malloc(sizeof(double) * (size_t)N * n_terms);
    double *pot = (double *)calloc((size_t)N * n_terms, sizeof(double));
    for (long j = 0; j < N; j++) {
        charges[j * n_terms + 0] = 1;
        charges[j * n_terms + 1] = xs[j];
        charges[j * n_terms + 2] = ys[j];
        charges[j * n_terms + 3] = xs[j] * xs[j] + ys[j] * ys[j];
    }
    int nb = orc_n_boxes_2d(min_coord, max_coord, ipi, min_num_intervals);
    int n1d = p * nb;
    int M = 2 * n1d;
    double *blo = (double *)malloc(sizeof(double) * 2 * (size_t)nb * nb);
    double *bhi = (double *)malloc(sizeof(double) * 2 * (size_t)nb * nb);
    double *sp = (double *)malloc(sizeof(double) * (size_t)p);
    double *xt = (double *)calloc((size_t)n1d, sizeof(double));
    double *yt = (double *)calloc((size_t)n1d, sizeof(double));
    double *kt = (double *)malloc(sizeof(double) * 2 * (size_t)M * (M / 2 + 1));
    /* argument order as at tsne.cpp:1092 (x_tilde and y_tilde swapped by the caller; identical grids) */
    orc_precompute_2d(max_coord, min_coord, max_coord, min_coord, nb, p, ORC_SQUARED_CAUCHY, blo, bhi, sp, xt, yt, kt, 1.0);
    orc_n_body_fft_2d(N, n_terms, xs, ys, charges, nb, p, blo, bhi, sp, kt, pot, nthreads);

    double sum_Q = 0;
    for (long i = 0; i < N; i++) {
        double phi1 = pot[i * n_terms + 0], phi2 = pot[i * n_terms + 1];
        double phi3 = pot[i * n_terms + 2], phi4 = pot[i * n_terms + 3];
        sum_Q += (1 + xs[i] * xs[i] + ys[i] * ys[i]) * phi1 - 2 * (xs[i] * phi2 + ys[i] * phi3) + phi4;
    }
    sum_Q -= N;
    *sum_Q_out = sum_Q;

    double *pos_f = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    orc_attractive(row_P, col_P, val_P, Y, N, 2, 1.0, nthreads, pos_f);
    for (long i = 0; i < N; i++) {
        double nx = (xs[i] * pot[i * n_terms] - pot[i * n_terms + 1]) / sum_Q;
        double ny = (ys[i] * pot[i * n_terms] - pot[i * n_terms + 2]) / sum_Q;
        dC[i * 2 + 0] = pos_f[i * 2] - nx;
        dC[i * 2 + 1] = pos_f[i * 2 + 1] - ny;
    }
    free(pos_f); free(pot); free(charges); free(xs); free(ys);
    free(blo); free(bhi); free(sp); free(xt); free(yt); free(kt);
}

/* ---- tsne.cpp:860-1024  computeFftGradientVariableDf (2D, df != 1) ---- */
void orc_grad_2d_df(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double df,
                    double *sum_Q_out) {
    for (long i = 0; i < (long)N * 2; i++) dC[i] = 0.0;
    double *xs = (double *)malloc(sizeof(double) * (size_t)N);
    double *ys = (double *)malloc(sizeof(double) * (size_t)N);
    for (long i = 0; i < N; i++) { xs[i] = Y[i * 2]; ys[i] = Y[i * 2 + 1]; }
    double min_coord, max_coord;
    orc_minmax_2d_quirk(Y, N, &min_coord, &max_coord);
    int nb = orc_n_boxes_2d(min_coord, max_coord, ipi, min_num_intervals);

    int sq_terms = 3;
    double *sqc = (double *)malloc(sizeof(double) * (size_t)N * sq_terms);
    double *sqp = (double *)calloc((size_t)N * sq_terms, sizeof(double));
    for (long j = 0; j < N; j++) {
        sqc[j * sq_terms + 0] = xs[j];
        sqc[j * sq_terms + 1] = ys[j];
        sqc[j * sq_terms + 2] = 1;
    }
    int n1d = p * nb, M = 2 * n1d;
    double *blo = (double *)malloc(sizeof(double) * 2 * (size_t)nb * nb);
    double *bhi = (double *)malloc(sizeof(double) * 2 * (size_t)nb * nb);
    double *sp = (double *)malloc(sizeof(double) * (size_t)p);
    double *xt = (double *)calloc((size_t)n1d, sizeof(double));
    double *yt = (double *)calloc((size_t)n1d, sizeof(double));
    double *kt = (double *)malloc(sizeof(double) * 2 * (size_t)M * (M / 2 + 1));
    orc_precompute_2d(max_coord, min_coord, max_coord, min_coord, nb, p, ORC_SQUARED_GENERAL, blo, bhi, sp, xt, yt, kt, df);
    orc_n_body_fft_2d(N, sq_terms, xs, ys, sqc, nb, p, blo, bhi, sp, kt, sqp, nthreads);

    double *nsc = (double *)malloc(sizeof(double) * (size_t)N);
    double *nsp = (double *)calloc((size_t)N, sizeof(double));
    for (long j = 0; j < N; j++) nsc[j] = 1;
    orc_precompute_2d(max_coord, min_coord, max_coord, min_coord, nb, p, ORC_GENERAL, blo, bhi, sp, xt, yt, kt, df);
    orc_n_body_fft_2d(N, 1, xs, ys, nsc, nb, p, blo, bhi, sp, kt, nsp, nthreads);

    double sum_Q = 0;
    for (long i = 0; i < N; i++) sum_Q += nsp[i];
    sum_Q -= N;

    double *pos_f = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    orc_attractive(row_P, col_P, val_P, Y, N, 2, df, nthreads, pos_f);
    for (long i = 0; i < N; i++) {
        double h2 = sqp[i * sq_terms], h3 = sqp[i * sq_terms + 1], h4 = sqp[i * sq_terms + 2];
        double nx = (xs[i] * h4 - h2) / sum_Q;
        double ny = (ys[i] * h4 - h3) / sum_Q;
        dC[i * 2 + 0] = (pos_f[i * 2] - nx);
        dC[i * 2 + 1] = (pos_f[i * 2 + 1] - ny);
    }
    *sum_Q_out = sum_Q;
    free(pos_f); free(sqc); free(sqp); free(nsc); free(nsp); free(xs); free(ys);
    free(blo); free(bhi); free(sp); free(xt); free(yt); free(kt);
}

static void minmax_1d(const double *Y, int N, double *mn, double *mx) {
    double a = INFINITY, b = -INFINITY;
    for (long i = 0; i < N; i++) {
        if (Y[i] < a) a = Y[i];
        if (Y[i] > b) b = Y[i];
    }
    *mn = a;
    *mx = b;
}

/* ---- tsne.cpp:758-856  computeFftGradientOneD (1D, df = 1) ---- */
void orc_grad_1d(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                 double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double *sum_Q_out) {
    for (long i = 0; i < N; i++) dC[i] = 0.0;
    double y_min, y_max;
    minmax_1d(Y, N, &y_min, &y_max);
    int nb = orc_n_boxes_1d(y_min, y_max, ipi, min_num_intervals);
    int n_terms = 3;
    double *ch = (double *)malloc(sizeof(double) * (size_t)N * n_terms);
    double *pot = (double *)calloc((size_t)N * n_terms, sizeof(double));
    for (long j = 0; j < N; j++) {
        ch[j * n_terms + 0] = 1;
        ch[j * n_terms + 1] = Y[j];
        ch[j * n_terms + 2] = Y[j] * Y[j];
    }
    double *blo = (double *)malloc(sizeof(double) * (size_t)nb);
    double *bhi = (double *)malloc(sizeof(double) * (size_t)nb);
    double *sp = (double *)malloc(sizeof(double) * (size_t)p);
    double *yt = (double *)calloc((size_t)p * nb, sizeof(double));
    double *kv = (double *)malloc(sizeof(double) * 2 * 2 * (size_t)p * nb);
    orc_precompute_1d(y_min, y_max, nb, p, ORC_SQUARED_CAUCHY, blo, bhi, sp, yt, kv, 1.0);
    orc_n_body_fft_1d(N, n_terms, Y, ch, nb, p, blo, bhi, sp, yt, kv, pot);
    double sum_Q = 0;
    for (long i = 0; i < N; i++) {
        double phi1 = pot[i * n_terms + 0], phi2 = pot[i * n_terms + 1], phi3 = pot[i * n_terms + 2];
        sum_Q += (1 + Y[i] * Y[i]) * phi1 - 2 * (Y[i] * phi2) + phi3;
    }
    sum_Q -= N;
    *sum_Q_out = sum_Q;
    double *pos_f = (double *)malloc(sizeof(double) * (size_t)N);
    orc_attractive(row_P, col_P, val_P, Y, N, 1, 1.0, nthreads, pos_f);
    for (long n = 0; n < N; n++) {
        double neg = (Y[n] * pot[n * n_terms] - pot[n * n_terms + 1]) / sum_Q;
        dC[n] = pos_f[n] - neg;
    }
    free(pos_f); free(ch); free(pot); free(blo); free(bhi); free(sp); free(yt); free(kv);
}

/* ---- tsne.cpp:647-754  computeFftGradientOneDVariableDf (1D, df != 1) ---- */
void orc_grad_1d_df(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double df,
                    double *sum_Q_out) {
    for (long i = 0; i < N; i++) dC[i] = 0.0;
    double y_min, y_max;
    minmax_1d(Y, N, &y_min, &y_max);
    int nb = orc_n_boxes_1d(y_min, y_max, ipi, min_num_intervals);
    int sq_terms = 2;
    double *sqc = (double *)malloc(sizeof(double) * (size_t)N * sq_terms);
    double *sqp = (double *)calloc((size_t)N * sq_terms, sizeof(double));
    for (long j = 0; j < N; j++) {
        sqc[j * sq_terms + 0] = Y[j];
        sqc[j * sq_terms + 1] = 1;
    }
    double *blo = (double *)malloc(sizeof(double) * (size_t)nb);
    double *bhi = (double *)malloc(sizeof(double) * (size_t)nb);
    double *sp = (double *)malloc(sizeof(double) * (size_t)p);
    double *yt = (double *)calloc((size_t)p * nb, sizeof(double));
    double *kv = (double *)malloc(sizeof(double) * 2 * 2 * (size_t)p * nb);
    orc_precompute_1d(y_min, y_max, nb, p, ORC_SQUARED_GENERAL, blo, bhi, sp, yt, kv, df);
    orc_n_body_fft_1d(N, sq_terms, Y, sqc, nb, p, blo, bhi, sp, yt, kv, sqp);
    double *nsc = (double *)malloc(sizeof(double) * (size_t)N);
    double *nsp = (double *)calloc((size_t)N, sizeof(double));
    for (long j = 0; j < N; j++) nsc[j] = 1;
    orc_precompute_1d(y_min, y_max, nb, p, ORC_GENERAL, blo, bhi, sp, yt, kv, df);
    orc_n_body_fft_1d(N, 1, Y, nsc, nb, p, blo, bhi, sp, yt, kv, nsp);
    double sum_Q = 0;
    for (long i = 0; i < N; i++) sum_Q += nsp[i];
    sum_Q -= N;
    *sum_Q_out = sum_Q;
    double *pos_f = (double *)malloc(sizeof(double) * (size_t)N);
    orc_attractive(row_P, col_P, val_P, Y, N, 1, df, nthreads, pos_f);
    for (long i = 0; i < N; i++) {
        double h2 = sqp[i * sq_terms], h4 = sqp[i * sq_terms + 1];
        double neg = (Y[i] * h4 - h2) / sum_Q;
        dC[i] = (pos_f[i] - neg);
    }
    free(pos_f); free(sqc); free(sqp); free(nsc); free(nsp); free(blo); free(bhi); free(sp); free(yt); free(kv);
}

/* tsne.cpp:444-464 */
int orc_compute_fft_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                             const double *Y, int N, int D, double *dC, int p, double ipi,
                             int min_num_intervals, unsigned nthreads, double df, double *sum_Q) {
    if (D == 1) {
        if (df == 1.0) orc_grad_1d(row_P, col_P, val_P, Y, N, dC, p, ipi, min_num_intervals, nthreads, sum_Q);
        else orc_grad_1d_df(row_P, col_P, val_P, Y, N, dC, p, ipi, min_num_intervals, nthreads, df, sum_Q);
        return 0;
    }
    if (D == 2) {
        if (df == 1.0) orc_grad_2d(row_P, col_P, val_P, Y, N, dC, p, ipi, min_num_intervals, nthreads, sum_Q);
        else orc_grad_2d_df(row_P, col_P, val_P, Y, N, dC, p, ipi, min_num_intervals, nthreads, df, sum_Q);
        return 0;
    }
    return -1;
}

/* tsne.cpp:1172-1228 (repulsion and Z of computeExactGradientTest) */
void orc_exact_repulsion(const double *Y, int N, int D, double df, double *neg_f, double *sum_Q_out) {
    double sum_Q = 0;
    for (long n = 0; n < N; n++)
        for (long m = 0; m < N; m++) {
            if (n == m) continue;
            double dd = 0;
            for (int d = 0; d < D; d++) dd += (Y[n * D + d] - Y[m * D + d]) * (Y[n * D + d] - Y[m * D + d]);
            sum_Q += 1.0 / pow(1.0 + dd / df, df);
        }
    for (long n = 0; n < N; n++) {
        double acc[2] = {0, 0};
        for (long m = 0; m < N; m++) {
            if (n == m) continue;
            double dd = 0;
            for (int d = 0; d < D; d++) dd += (Y[n * D + d] - Y[m * D + d]) * (Y[n * D + d] - Y[m * D + d]);
            double q = 1.0 / pow(1.0 + dd / df, df);
            double w = pow(q, (df + 1.0) / df);
            for (int d = 0; d < D; d++) acc[d] += w * (Y[n * D + d] - Y[m * D + d]) / sum_Q;
        }
        for (int d = 0; d < D; d++) neg_f[n * D + d] = acc[d];
    }
    *sum_Q_out = sum_Q;
}

/* ---- optimiser ---- */

/* tsne.cpp:1851-1876 */
void orc_zero_mean(double *X, int N, int D) {
    double *mean = (double *)calloc((size_t)D, sizeof(double));
    long nD = 0;
    for (int n = 0; n < N; n++) {
        for (int d = 0; d < D; d++) mean[d] += X[nD + d];
        nD += D;
    }
    for (int d = 0; d < D; d++) mean[d] /= (double)N;
    nD = 0;
    for (int n = 0; n < N; n++) {
        for (int d = 0; d < D; d++) X[nD + d] -= mean[d];
        nD += D;
    }
    free(mean);
}

/* tsne.h:36 */
static inline double sgn(double x) { return (x == .0 ? .0 : (x < .0 ? -1.0 : 1.0)); }

/* tsne.cpp:478-514.  plain_descent selects the no-momentum branch (tsne.cpp:487-490). */
void orc_update(double *Y, double *uY, double *gains, const double *dY, int N, int D, double momentum,
                double learning_rate, double max_step_norm, int plain_descent) {
    long ND = (long)N * D;
    if (plain_descent) {
        for (long i = 0; i < ND; i++) Y[i] = Y[i] - dY[i];
        return;
    }
    for (long i = 0; i < ND; i++) gains[i] = (sgn(dY[i]) != sgn(uY[i])) ? (gains[i] + .2) : (gains[i] * .8);
    for (long i = 0; i < ND; i++) if (gains[i] < .01) gains[i] = .01;
    for (long i = 0; i < ND; i++) uY[i] = momentum * uY[i] - learning_rate * gains[i] * dY[i];
    if (max_step_norm > 0) {
        for (long i = 0; i < N; i++) {
            double step = 0;
            for (int j = 0; j < D; j++) step += uY[i * D + j] * uY[i * D + j];
            step = sqrt(step);
            if (step > max_step_norm)
                for (int j = 0; j < D; j++) uY[i * D + j] *= (max_step_norm / step);
        }
    }
    for (long i = 0; i < ND; i++) Y[i] = Y[i] + uY[i];
}

/* tsne.cpp:1329-1355 with the `C += temp` race removed: per-row partials summed in row order. */
double orc_evaluate_error_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                              const double *Y, int N, int D, double sum_Q, double df) {
    double C = .0;
    double buff[2];
    for (int n = 0; n < N; n++) {
        long ind1 = (long)n * D;
        double temp = 0;
        for (uint32_t i = row_P[n]; i < row_P[n + 1]; i++) {
            double Q = .0;
            long ind2 = (long)col_P[i] * D;
            for (int d = 0; d < D; d++) buff[d] = Y[ind1 + d];
            for (int d = 0; d < D; d++) buff[d] -= Y[ind2 + d];
            for (int d = 0; d < D; d++) Q += buff[d] * buff[d];
            Q = pow(1.0 / (1.0 + Q / df), df) / sum_Q;
            temp += val_P[i] * log((val_P[i] + FLT_MIN) / (Q + FLT_MIN));
        }
        C += temp;
    }
    return C;
}

/* TSNE::run's hot loop (tsne.cpp:389-412, :437-577) around one of its three gradient sources
 * (tsne.cpp:440-467): the FFT variants (nbody_algorithm 2), Barnes-Hut (1) or the exact dense gradient
 * (theta = 0), each with its own KL (tsne.cpp:551-559).  P (nval entries: the CSR values, or the dense
 * N x N matrix) is a private copy exaggerated in place as the reference does. */
enum { RUN_FFT = 0, RUN_BH = 1, RUN_EXACT = 2 };

static int run_loop(int kind, const uint32_t *row_P, const uint32_t *col_P, const double *P_in, long nval,
                    int N, int no_dims, double *Y, const orc_run_params *prm, double theta, double *costs) {
    long ND = (long)N * no_dims;
    double early = prm->early_exag_coeff;
    /* the automatic coefficient sums CSR rows (tsne.cpp:391-402); the exact mode has no CSR there */
    if (early == 0 && kind == RUN_EXACT) return -1;
    double *P = (double *)malloc(sizeof(double) * (size_t)(nval > 0 ? nval : 1));
    double *dY = (double *)malloc(sizeof(double) * (size_t)(ND > 0 ? ND : 1));
    double *uY = (double *)calloc((size_t)(ND > 0 ? ND : 1), sizeof(double));
    double *gains = (double *)malloc(sizeof(double) * (size_t)(ND > 0 ? ND : 1));
    memcpy(P, P_in, sizeof(double) * (size_t)nval);
    for (long i = 0; i < ND; i++) gains[i] = 1.0;
    double momentum = prm->momentum;
    if (early == 0) {
        double max_sum_cols = 0;
        for (int n = 0; n < N; n++) {
            double rs = 0;
            for (uint32_t i = row_P[n]; i < row_P[n + 1]; i++) rs += P[i];
            if (rs > max_sum_cols) max_sum_cols = rs;
        }
        early = (1.0 / (prm->learning_rate * max_sum_cols));
    }
    for (long i = 0; i < nval; i++) P[i] *= early;
    for (int i = 0; i < prm->max_iter; i++) costs[i] = 0.0;

    double sum_Q = 0;
    int rc = 0;
    for (int iter = 0; iter < prm->max_iter; iter++) {
        if (kind == RUN_FFT)
            rc = orc_compute_fft_gradient(row_P, col_P, P, Y, N, no_dims, dY, prm->n_interpolation_points,
                                          prm->intervals_per_integer, prm->min_num_intervals, prm->nthreads,
                                          prm->df, &sum_Q);
        else if (kind == RUN_BH)
            orc_bh_gradient(row_P, col_P, P, Y, N, no_dims, theta, dY, &sum_Q);
        else
            orc_exact_gradient(P, Y, N, no_dims, prm->df, dY, &sum_Q);
        if (rc) break;
        if (prm->no_momentum_during_exag) {
            if (iter > prm->stop_lying_iter)
                orc_update(Y, uY, gains, dY, N, no_dims, momentum, prm->learning_rate, -1.0, 0);
            else
                orc_update(Y, uY, gains, dY, N, no_dims, momentum, prm->learning_rate, -1.0, 1);
        } else {
            orc_update(Y, uY, gains, dY, N, no_dims, momentum, prm->learning_rate, prm->max_step_norm, 0);
        }
        orc_zero_mean(Y, N, no_dims);
        if (iter == prm->stop_lying_iter)
            for (long i = 0; i < nval; i++) P[i] /= early;
        if (iter == prm->start_late_exag_iter)
            for (long i = 0; i < nval; i++) P[i] *= prm->late_exag_coeff;
        if (iter == prm->mom_switch_iter) momentum = prm->final_momentum;
        if ((iter + 1) % 50 == 0 || iter == prm->max_iter - 1) {
            double C;
            if (kind == RUN_FFT) C = orc_evaluate_error_fft(row_P, col_P, P, Y, N, no_dims, sum_Q, prm->df);
            else if (kind == RUN_BH) C = orc_bh_error(row_P, col_P, P, Y, N, no_dims, theta);
            else C = orc_evaluate_error_exact(P, Y, N, no_dims, prm->df);
            if (iter < prm->stop_lying_iter && prm->stop_lying_iter != -1) C = C / early - log(early);
            if (iter >= prm->start_late_exag_iter && prm->start_late_exag_iter != -1)
                C = C / prm->late_exag_coeff - log(prm->late_exag_coeff);
            costs[iter] = C;
        }
    }
    free(P); free(dY); free(uY); free(gains);
    return rc;
}

int orc_run_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N, int no_dims,
                double *Y, const orc_run_params *prm, double *costs) {
    return run_loop(RUN_FFT, row_P, col_P, val_P, (long)row_P[N], N, no_dims, Y, prm, 0.0, costs);
}

int orc_run_bh(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N, int no_dims,
               double *Y, const orc_run_params *prm, double theta, double *costs) {
    return run_loop(RUN_BH, row_P, col_P, val_P, (long)row_P[N], N, no_dims, Y, prm, theta, costs);
}

int orc_run_exact(const double *P, int N, int no_dims, double *Y, const orc_run_params *prm, double *costs) {
    return run_loop(RUN_EXACT, NULL, NULL, P, (long)N * N, N, no_dims, Y, prm, 0.0, costs);
}

/* ---- affinities ---- */

/* tsne.cpp:1394-1469 with the starting beta exposed (the reference uses 1.0 at :1415). */
double orc_d2s_bisect(const double *D, double *P, int N, double perplexity, double beta0, int ifSquared,
                      int *iters, double *H_out) {
    double min_beta = -DBL_MAX, max_beta = DBL_MAX, tol = 1e-5;
    int max_iter = 200, iter = 0, found = 0;
    double beta = beta0, sum_P = DBL_MIN, H = 0.0;
    while (!found && iter < max_iter) {
        for (int m = 0; m < N; m++) P[m] = exp(-beta * (ifSquared ? D[m] : D[m] * D[m]));
        sum_P = DBL_MIN;
        for (int m = 0; m < N; m++) sum_P += P[m];
        H = 0.0;
        for (int m = 0; m < N; m++) H += beta * ((ifSquared ? D[m] : D[m] * D[m]) * P[m]);
        H = (H / sum_P) + log(sum_P);
        double Hdiff = H - log(perplexity);
        if (Hdiff < tol && -Hdiff < tol) {
            found = 1;
        } else {
            if (Hdiff > 0) {
                min_beta = beta;
                if (max_beta == DBL_MAX || max_beta == -DBL_MAX) beta *= 2.0;
                else beta = (beta + max_beta) / 2.0;
            } else {
                max_beta = beta;
                if (min_beta == -DBL_MAX || min_beta == DBL_MAX) beta /= 2.0;
                else beta = (beta + min_beta) / 2.0;
            }
        }
        iter++;
    }
    for (int m = 0; m < N; m++) P[m] /= sum_P;
    if (iters) *iters = iter;
    if (H_out) *H_out = H;
    return beta;
}

/* tsne.cpp:1394-1469 */
double orc_distances2similarities(const double *D, double *P, int N, int n, double perplexity,
                                  double sigma, int ifSquared) {
    double sum_P, beta;
    if (perplexity > 0) {
        double min_beta = -DBL_MAX, max_beta = DBL_MAX, tol = 1e-5;
        int max_iter = 200, iter = 0, found = 0;
        beta = 1.0;
        sum_P = DBL_MIN;
        while (!found && iter < max_iter) {
            for (int m = 0; m < N; m++) P[m] = exp(-beta * (ifSquared ? D[m] : D[m] * D[m]));
            if (n >= 0) P[n] = DBL_MIN;
            sum_P = DBL_MIN;
            for (int m = 0; m < N; m++) sum_P += P[m];
            double H = 0.0;
            for (int m = 0; m < N; m++) H += beta * ((ifSquared ? D[m] : D[m] * D[m]) * P[m]);
            H = (H / sum_P) + log(sum_P);
            double Hdiff = H - log(perplexity);
            if (Hdiff < tol && -Hdiff < tol) {
                found = 1;
            } else {
                if (Hdiff > 0) {
                    min_beta = beta;
                    if (max_beta == DBL_MAX || max_beta == -DBL_MAX) beta *= 2.0;
                    else beta = (beta + max_beta) / 2.0;
                } else {
                    max_beta = beta;
                    if (min_beta == -DBL_MAX || min_beta == DBL_MAX) beta /= 2.0;
                    else beta = (beta + min_beta) / 2.0;
                }
            }
            iter++;
        }
    } else {
        beta = 1 / (2 * sigma * sigma);
        for (int m = 0; m < N; m++) P[m] = exp(-beta * (ifSquared ? D[m] : D[m] * D[m]));
        if (n >= 0) P[n] = DBL_MIN;
        sum_P = DBL_MIN;
        for (int m = 0; m < N; m++) sum_P += P[m];
    }
    for (int m = 0; m < N; m++) P[m] /= sum_P;
    return beta;
}

/* tsne.cpp:1474-1500 */
double orc_distances2similarities_list(const double *D, double *P, int N, int n, double perplexity,
                                       double sigma, int ifSquared, int perplexity_list_length,
                                       const double *perplexity_list) {
    (void)ifSquared;
    if (perplexity != 0) return orc_distances2similarities(D, P, N, n, perplexity, sigma, 1);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)N);
    double beta = orc_distances2similarities(D, P, N, n, perplexity_list[0], sigma, 1);
    for (int m = 1; m < perplexity_list_length; m++) {
        beta = orc_distances2similarities(D, tmp, N, n, perplexity_list[m], sigma, 1);
        for (int i = 0; i < N; i++) P[i] += tmp[i];
    }
    for (int i = 0; i < N; i++) P[i] /= perplexity_list_length;
    free(tmp);
    return beta;
}

/* Exact kNN by brute force; distances as vptree.h's euclidean_distance (sqrt of the ordered sum). */
typedef struct {
    const double *X;
    int N, D, K;
    int32_t *idx;
    double *dist;
} knn_ctx;

static void knn_body(int lo, int hi, void *c_) {
    knn_ctx *c = (knn_ctx *)c_;
    int K = c->K, D = c->D, N = c->N;
    double *bd = (double *)malloc(sizeof(double) * (size_t)(K + 1));
    int32_t *bi = (int32_t *)malloc(sizeof(int32_t) * (size_t)(K + 1));
    for (int n = lo; n < hi; n++) {
        int cnt = 0;
        const double *xn = c->X + (long)n * D;
        for (int m = 0; m < N; m++) {
            if (m == n) continue;
            const double *xm = c->X + (long)m * D;
            double dd = .0;
            for (int d = 0; d < D; d++) { double t = xn[d] - xm[d]; dd += t * t; }
            double dist = sqrt(dd);
            if (cnt == K && dist >= bd[K - 1]) continue;
            int pos = cnt < K ? cnt : K - 1;
            while (pos > 0 && bd[pos - 1] > dist) { bd[pos] = bd[pos - 1]; bi[pos] = bi[pos - 1]; pos--; }
            bd[pos] = dist;
            bi[pos] = m;
            if (cnt < K) cnt++;
        }
        for (int k = 0; k < K; k++) {
            c->idx[(long)n * K + k] = bi[k];
            c->dist[(long)n * K + k] = bd[k];
        }
    }
    free(bd);
    free(bi);
}

void orc_knn_exact(const double *X, int N, int D, int K, int32_t *idx, double *dist, unsigned nthreads) {
    knn_ctx c = {X, N, D, K, idx, dist};
    orc_parallel_for(nthreads, N, knn_body, &c);
}

/* tsne.cpp:1684-1706: per point, distances2similarities on the K non-squared distances
 * (ifSquared = false, self excluded), then the row is stored in kNN order. */
void orc_rows_from_knn(const int32_t *idx, const double *dist, int N, int K, double perplexity,
                       double sigma, int perplexity_list_length, const double *perplexity_list,
                       uint32_t *row_P, uint32_t *col_P, double *val_P) {
    row_P[0] = 0;
    for (int n = 0; n < N; n++) row_P[n + 1] = row_P[n] + (uint32_t)K;
    double *cur = (double *)malloc(sizeof(double) * (size_t)K);
    for (int n = 0; n < N; n++) {
        /* The call site passes ifSquared = false, but the list overload forwards ifSquared = true
         * (tsne.cpp:1479, :1486, :1489), so the non-squared kNN distances enter exp(-beta * d)
         * unsquared.  Restated as the reference behaves. */
        orc_distances2similarities_list(dist + (long)n * K, cur, K, -1, perplexity, sigma, 0,
                                        perplexity_list_length, perplexity_list);
        for (int m = 0; m < K; m++) {
            col_P[row_P[n] + m] = (uint32_t)idx[(long)n * K + m];
            val_P[row_P[n] + m] = cur[m];
        }
    }
    free(cur);
}

/* tsne.cpp:1730-1828 */
int orc_symmetrize(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N,
                   uint32_t *sym_row_P, uint32_t *sym_col_P, double *sym_val_P) {
    int *row_counts = (int *)calloc((size_t)N, sizeof(int));
    for (int n = 0; n < N; n++) {
        for (uint32_t i = row_P[n]; i < row_P[n + 1]; i++) {
            int present = 0;
            for (uint32_t m = row_P[col_P[i]]; m < row_P[col_P[i] + 1]; m++)
                if (col_P[m] == (uint32_t)n) present = 1;
            if (present) row_counts[n]++;
            else { row_counts[n]++; row_counts[col_P[i]]++; }
        }
    }
    int no_elem = 0;
    for (int n = 0; n < N; n++) no_elem += row_counts[n];
    sym_row_P[0] = 0;
    for (int n = 0; n < N; n++) sym_row_P[n + 1] = sym_row_P[n] + (uint32_t)row_counts[n];
    int *offset = (int *)calloc((size_t)N, sizeof(int));
    for (int n = 0; n < N; n++) {
        for (uint32_t i = row_P[n]; i < row_P[n + 1]; i++) {
            int present = 0;
            uint32_t c = col_P[i];
            for (uint32_t m = row_P[c]; m < row_P[c + 1]; m++) {
                if (col_P[m] == (uint32_t)n) {
                    present = 1;
                    if ((uint32_t)n <= c) {
                        sym_col_P[sym_row_P[n] + offset[n]] = c;
                        sym_col_P[sym_row_P[c] + offset[c]] = (uint32_t)n;
                        sym_val_P[sym_row_P[n] + offset[n]] = val_P[i] + val_P[m];
                        sym_val_P[sym_row_P[c] + offset[c]] = val_P[i] + val_P[m];
                    }
                }
            }
            if (!present) {
                sym_col_P[sym_row_P[n] + offset[n]] = c;
                sym_col_P[sym_row_P[c] + offset[c]] = (uint32_t)n;
                sym_val_P[sym_row_P[n] + offset[n]] = val_P[i];
                sym_val_P[sym_row_P[c] + offset[c]] = val_P[i];
            }
            if (!present || ((uint32_t)n <= c)) {
                offset[n]++;
                if (c != (uint32_t)n) offset[c]++;
            }
        }
    }
    for (int i = 0; i < no_elem; i++) sym_val_P[i] /= 2.0;
    free(offset);
    free(row_counts);
    return no_elem;
}

/* ---- exact t-SNE (theta = 0): dense P ---- */

/* tsne.cpp:1832-1848: the upper triangle in order d = 0..D-1 from 0.0, mirrored; zero diagonal. */
void orc_squared_distances(const double *X, int N, int D, double *DD) {
    for (long n = 0; n < N; n++) {
        DD[n * N + n] = 0.0;
        for (long m = n + 1; m < N; m++) {
            double acc = 0.0;
            for (int d = 0; d < D; d++) acc += (X[n * D + d] - X[m * D + d]) * (X[n * D + d] - X[m * D + d]);
            DD[n * N + m] = acc;
            DD[m * N + n] = acc;
        }
    }
}

typedef struct {
    const double *DD;
    double *P;
    int N, L;
    double perplexity, sigma;
    const double *list;
} gpd_ctx;

static void gpd_body(int lo, int hi, void *c_) {
    gpd_ctx *c = (gpd_ctx *)c_;
    for (int n = lo; n < hi; n++)
        orc_distances2similarities_list(c->DD + (long)n * c->N, c->P + (long)n * c->N, c->N, n, c->perplexity,
                                        c->sigma, 1, c->L, c->list);
}

/* tsne.cpp:1503-1530 (computeGaussianPerplexity, dense overload): every row of the squared-distance
 * matrix through distances2similarities with the row's own index as the self entry; the list overload
 * is always called with ifSquared = true. */
void orc_gaussian_perplexity_dense(const double *X, int N, int D, double perplexity, double sigma,
                                   int perplexity_list_length, const double *perplexity_list, double *P,
                                   unsigned nthreads) {
    double *DD = (double *)malloc(sizeof(double) * (size_t)N * N);
    orc_squared_distances(X, N, D, DD);
    gpd_ctx c = {DD, P, N, perplexity_list_length, perplexity, sigma, perplexity_list};
    orc_parallel_for(nthreads, N, gpd_body, &c);
    free(DD);
}

/* tsne.cpp:196-216: P[n][m] = P[m][n] = P[n][m] + P[m][n] for n < m (diagonal untouched), then every
 * entry divided by the total (summed in row-major order).  Returns the total. */
double orc_exact_symmetrize_normalize(double *P, int N) {
    for (long n = 0; n < N; n++)
        for (long m = n + 1; m < N; m++) {
            P[n * N + m] += P[m * N + n];
            P[m * N + n] = P[n * N + m];
        }
    double sum_P = .0;
    const long NN = (long)N * N;
    for (long i = 0; i < NN; i++) sum_P += P[i];
    for (long i = 0; i < NN; i++) P[i] /= sum_P;
    return sum_P;
}

/* tsne.cpp:1232-1282: Q = 1/(1 + d2/df), Qpow = Q^df, sum_Q over n != m, and
 * dC[n] = sum_{m != n} (Y[n] - Y[m]) (P[n][m] - Qpow/sum_Q) Q (no factor 4). */
void orc_exact_gradient(const double *P, const double *Y, int N, int D, double df, double *dC, double *sum_Q_out) {
    const long NN = (long)N * N;
    double *DD = (double *)malloc(sizeof(double) * (size_t)NN);
    double *Q = (double *)malloc(sizeof(double) * (size_t)NN);
    double *Qpow = (double *)malloc(sizeof(double) * (size_t)NN);
    for (long i = 0; i < (long)N * D; i++) dC[i] = 0.0;
    orc_squared_distances(Y, N, D, DD);
    double sum_Q = .0;
    for (long n = 0; n < N; n++)
        for (long m = 0; m < N; m++)
            if (n != m) {
                Q[n * N + m] = 1.0 / (1.0 + DD[n * N + m] / df);
                Qpow[n * N + m] = pow(Q[n * N + m], df);
                sum_Q += Qpow[n * N + m];
            }
    for (long n = 0; n < N; n++)
        for (long m = 0; m < N; m++)
            if (n != m) {
                double mult = (P[n * N + m] - (Qpow[n * N + m] / sum_Q)) * Q[n * N + m];
                for (int d = 0; d < D; d++) dC[n * D + d] += (Y[n * D + d] - Y[m * D + d]) * mult;
            }
    if (sum_Q_out) *sum_Q_out = sum_Q;
    free(DD);
    free(Q);
    free(Qpow);
}

/* tsne.cpp:1286-1326: Q normalised by its own sum (started at DBL_MIN) with DBL_MIN on the diagonal;
 * C = sum over all N^2 entries of P log((P + FLT_MIN) / (Q + FLT_MIN)). */
double orc_evaluate_error_exact(const double *P, const double *Y, int N, int D, double df) {
    const long NN = (long)N * N;
    double *DD = (double *)malloc(sizeof(double) * (size_t)NN);
    double *Q = (double *)malloc(sizeof(double) * (size_t)NN);
    orc_squared_distances(Y, N, D, DD);
    double sum_Q = DBL_MIN;
    for (long n = 0; n < N; n++)
        for (long m = 0; m < N; m++) {
            if (n != m) {
                Q[n * N + m] = 1.0 / (1.0 + DD[n * N + m] / df);
                Q[n * N + m] = pow(Q[n * N + m], df);
                sum_Q += Q[n * N + m];
            } else {
                Q[n * N + m] = DBL_MIN;
            }
        }
    for (long i = 0; i < NN; i++) Q[i] /= sum_Q;
    double C = .0;
    for (long i = 0; i < NN; i++) C += P[i] * log((P[i] + FLT_MIN) / (Q[i] + FLT_MIN));
    free(DD);
    free(Q);
    return C;
}

/* ---- Barnes-Hut (nbody_algo = 1): SPTree (sptree.cpp) and computeGradient / evaluateError ---- */

/* A flattened SPTree: node k has center c[k*D+d] and half-width w[k*D+d] (the reference's Cell
 * "corner" and "width"), the online center of mass m[k*D+d], cum_size, at most one stored point, and
 * 2^D children allocated together at first_child[k] (or -1). */
typedef struct {
    int D, nc;
    long n, cap;
    double *c, *w, *m;
    unsigned *cum, *size, *pt;
    int *leaf;
    long *first_child;
    const double *Y;
} bh_tree;

static long bh_new_node(bh_tree *t, const double *center, const double *hw) {
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->c = (double *)realloc(t->c, sizeof(double) * t->cap * t->D);
        t->w = (double *)realloc(t->w, sizeof(double) * t->cap * t->D);
        t->m = (double *)realloc(t->m, sizeof(double) * t->cap * t->D);
        t->cum = (unsigned *)realloc(t->cum, sizeof(unsigned) * t->cap);
        t->size = (unsigned *)realloc(t->size, sizeof(unsigned) * t->cap);
        t->pt = (unsigned *)realloc(t->pt, sizeof(unsigned) * t->cap);
        t->leaf = (int *)realloc(t->leaf, sizeof(int) * t->cap);
        t->first_child = (long *)realloc(t->first_child, sizeof(long) * t->cap);
    }
    long k = t->n++;
    for (int d = 0; d < t->D; d++) {
        t->c[k * t->D + d] = center[d];
        t->w[k * t->D + d] = hw[d];
        t->m[k * t->D + d] = .0;
    }
    t->cum[k] = 0;
    t->size[k] = 0;
    t->pt[k] = 0;
    t->leaf[k] = 1;
    t->first_child[k] = -1;
    return k;
}

static int bh_insert(bh_tree *t, long k, unsigned i);

/* sptree.cpp:235-262 */
static void bh_subdivide(bh_tree *t, long k) {
    const int D = t->D;
    double nc[8], nw[8];
    long first = -1;
    for (int ch = 0; ch < t->nc; ch++) {
        unsigned div = 1;
        for (int d = 0; d < D; d++) {
            nw[d] = .5 * t->w[k * D + d];
            if ((ch / div) % 2 == 1) nc[d] = t->c[k * D + d] - .5 * t->w[k * D + d];
            else nc[d] = t->c[k * D + d] + .5 * t->w[k * D + d];
            div *= 2;
        }
        long id = bh_new_node(t, nc, nw);
        if (ch == 0) first = id;
    }
    t->first_child[k] = first;
    for (unsigned s = 0; s < t->size[k]; s++) {
        int ok = 0;
        for (int ch = 0; ch < t->nc; ch++)
            if (!ok) ok = bh_insert(t, first + ch, t->pt[k]);
    }
    t->size[k] = 0;
    t->leaf[k] = 0;
}

/* sptree.cpp:189-232 */
static int bh_insert(bh_tree *t, long k, unsigned i) {
    const int D = t->D;
    const double *p = t->Y + (long)i * D;
    for (int d = 0; d < D; d++) {
        if (t->c[k * D + d] - t->w[k * D + d] > p[d]) return 0;
        if (t->c[k * D + d] + t->w[k * D + d] < p[d]) return 0;
    }
    t->cum[k]++;
    double mult1 = (double)(t->cum[k] - 1) / (double)t->cum[k];
    double mult2 = 1.0 / (double)t->cum[k];
    for (int d = 0; d < D; d++) t->m[k * D + d] *= mult1;
    for (int d = 0; d < D; d++) t->m[k * D + d] += mult2 * p[d];
    if (t->leaf[k] && t->size[k] < 1) {
        t->pt[k] = i;
        t->size[k] = 1;
        return 1;
    }
    int dup = 0;
    for (unsigned s = 0; s < t->size[k]; s++) {
        int same = 1;
        for (int d = 0; d < D; d++)
            if (p[d] != t->Y[(long)t->pt[k] * D + d]) {
                same = 0;
                break;
            }
        dup = dup | same;
    }
    if (dup) return 1;
    if (t->leaf[k]) bh_subdivide(t, k);
    for (int ch = 0; ch < t->nc; ch++)
        if (bh_insert(t, t->first_child[k] + ch, i)) return 1;
    return 0;
}

/* sptree.cpp:57-84 (root cell: the mean, half-widths max(max - mean, mean - min) + 1e-5) and :277 */
static void bh_build(bh_tree *t, const double *Y, int N, int D) {
    memset(t, 0, sizeof(*t));
    t->D = D;
    t->nc = 1 << D;
    t->Y = Y;
    double mean[8] = {0}, mn[8], mx[8], w[8];
    for (int d = 0; d < D; d++) {
        mn[d] = DBL_MAX;
        mx[d] = -DBL_MAX;
    }
    for (long n = 0; n < N; n++)
        for (int d = 0; d < D; d++) {
            mean[d] += Y[n * D + d];
            if (Y[n * D + d] < mn[d]) mn[d] = Y[n * D + d];
            if (Y[n * D + d] > mx[d]) mx[d] = Y[n * D + d];
        }
    for (int d = 0; d < D; d++) mean[d] /= (double)N;
    for (int d = 0; d < D; d++) w[d] = fmax(mx[d] - mean[d], mean[d] - mn[d]) + 1e-5;
    bh_new_node(t, mean, w);
    for (unsigned i = 0; i < (unsigned)N; i++) bh_insert(t, 0, i);
}

static void bh_free(bh_tree *t) {
    free(t->c);
    free(t->w);
    free(t->m);
    free(t->cum);
    free(t->size);
    free(t->pt);
    free(t->leaf);
    free(t->first_child);
}

/* sptree.cpp:322-354 */
static void bh_nonedge(const bh_tree *t, long k, unsigned i, double theta, double *neg_f, double *sum_Q) {
    const int D = t->D;
    if (t->cum[k] == 0 || (t->leaf[k] && t->size[k] == 1 && t->pt[k] == i)) return;
    double Dd = .0;
    const double *y = t->Y + (long)i * D;
    for (int d = 0; d < D; d++) Dd += (y[d] - t->m[k * D + d]) * (y[d] - t->m[k * D + d]);
    double max_width = 0.0, cur;
    for (int d = 0; d < D; d++) {
        cur = t->w[k * D + d];
        max_width = (max_width > cur) ? max_width : cur;
    }
    if (t->leaf[k] || max_width / sqrt(Dd) < theta) {
        Dd = 1.0 / (1.0 + Dd);
        double mult = t->cum[k] * Dd;
        *sum_Q += mult;
        mult *= Dd;
        for (int d = 0; d < D; d++) neg_f[d] += mult * (y[d] - t->m[k * D + d]);
    } else {
        for (int ch = 0; ch < t->nc; ch++) bh_nonedge(t, t->first_child[k] + ch, i, theta, neg_f, sum_Q);
    }
}

/* tsne.cpp:596-640 with sptree.cpp:358-376 (edge forces) and nthreads = 1 */
void orc_bh_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                     int D, double theta, double *dC, double *sum_Q_out) {
    bh_tree t;
    bh_build(&t, Y, N, D);
    double *pos_f = (double *)calloc((size_t)N * D, sizeof(double));
    double *neg_f = (double *)calloc((size_t)N * D, sizeof(double));
    double *Q = (double *)calloc((size_t)N, sizeof(double));
    for (long n = 0; n < N; n++) {
        for (uint32_t e = row_P[n]; e < row_P[n + 1]; e++) {
            double Dd = 1.0;
            long m = col_P[e];
            for (int d = 0; d < D; d++) Dd += (Y[n * D + d] - Y[m * D + d]) * (Y[n * D + d] - Y[m * D + d]);
            Dd = val_P[e] / Dd;
            for (int d = 0; d < D; d++) pos_f[n * D + d] += Dd * (Y[n * D + d] - Y[m * D + d]);
        }
    }
    for (unsigned n = 0; n < (unsigned)N; n++) bh_nonedge(&t, 0, n, theta, neg_f + (long)n * D, Q + n);
    double sum_Q = .0;
    for (long n = 0; n < N; n++) sum_Q += Q[n];
    for (long e = 0; e < (long)N * D; e++) dC[e] = pos_f[e] - (neg_f[e] / sum_Q);
    if (sum_Q_out) *sum_Q_out = sum_Q;
    free(pos_f);
    free(neg_f);
    free(Q);
    bh_free(&t);
}

/* tsne.cpp:1359-1390 with nthreads = 1: sum_Q as one running sum over all points, then the KL over
 * the edges, rows in order. */
double orc_bh_error(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    int D, double theta) {
    bh_tree t;
    bh_build(&t, Y, N, D);
    double buff[8] = {0};
    double sum_Q = .0;
    for (unsigned n = 0; n < (unsigned)N; n++) bh_nonedge(&t, 0, n, theta, buff, &sum_Q);
    double C = .0;
    for (long n = 0; n < N; n++) {
        double temp = 0;
        for (uint32_t e = row_P[n]; e < row_P[n + 1]; e++) {
            double Qv = .0;
            long m = col_P[e];
            for (int d = 0; d < D; d++) buff[d] = Y[n * D + d];
            for (int d = 0; d < D; d++) buff[d] -= Y[m * D + d];
            for (int d = 0; d < D; d++) Qv += buff[d] * buff[d];
            Qv = (1.0 / (1.0 + Qv)) / sum_Q;
            temp += val_P[e] * log((val_P[e] + FLT_MIN) / (Qv + FLT_MIN));
        }
        C += temp;
    }
    bh_free(&t);
    return C;
}
