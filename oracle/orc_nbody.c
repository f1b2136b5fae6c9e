/*
 * orc_nbody.c -- restatement of src/nbodyfft.cpp (FIt-SNE 1.2.1) and the output kernels of
 * src/tsne.cpp:69-94.  TEST INFRASTRUCTURE: see fitsne_oracle.h.
 */
#include "fitsne_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* tsne.cpp:69-79 */
double orc_kernel_1d(int kind, double x, double y, double df) {
    switch (kind) {
    case ORC_SQUARED_CAUCHY: return pow(1.0 + pow(x - y, 2), -2);
    case ORC_GENERAL: return pow(1.0 + ((x - y) * (x - y)) / df, -(df));
    default: return pow(1.0 + ((x - y) * (x - y)) / df, -(df + 1.0));
    }
}

/* tsne.cpp:84-94 */
double orc_kernel_2d(int kind, double x1, double x2, double y1, double y2, double df) {
    switch (kind) {
    case ORC_SQUARED_CAUCHY: return pow(1.0 + pow(x1 - y1, 2) + pow(x2 - y2, 2), -2);
    case ORC_GENERAL: return pow(1.0 + ((x1 - y1) * (x1 - y1) + (x2 - y2) * (x2 - y2)) / df, -(df));
    default: return pow(1.0 + ((x1 - y1) * (x1 - y1) + (x2 - y2) * (x2 - y2)) / df, -(df + 1.0));
    }
}

/* ---- parallel_for.h:26-41 ---- */
typedef struct {
    orc_body_fn fn;
    void *ctx;
    int lo, hi;
} pf_arg;

static void *pf_tramp(void *a_) {
    pf_arg *a = (pf_arg *)a_;
    a->fn(a->lo, a->hi, a->ctx);
    return NULL;
}

void orc_parallel_for(unsigned nthreads, int n, orc_body_fn fn, void *ctx) {
    if (nthreads > 1) {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
        pf_arg *args = (pf_arg *)malloc(sizeof(pf_arg) * nthreads);
        for (unsigned t = 0; t < nthreads; t++) {
            args[t].fn = fn;
            args[t].ctx = ctx;
            args[t].lo = (int)((long)t * n / nthreads);
            args[t].hi = (t + 1) == nthreads ? n : (int)((long)(t + 1) * n / nthreads);
            pthread_create(&th[t], NULL, pf_tramp, &args[t]);
        }
        for (unsigned t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(th);
        free(args);
    } else {
        fn(0, n, ctx);
    }
}

/* ---- nbodyfft.cpp:9-70 ---- */
void orc_precompute_2d(double x_max, double x_min, double y_max, double y_min, int n_boxes,
                       int n_interpolation_points, int kernel, double *box_lower_bounds,
                       double *box_upper_bounds, double *y_tilde_spacings, double *y_tilde,
                       double *x_tilde, double *fft_kernel_tilde, double df) {
    (void)y_max;
    int n_total_boxes = n_boxes * n_boxes;
    double box_width = (x_max - x_min) / (double)n_boxes;
    for (int i = 0; i < n_boxes; i++) {
        for (int j = 0; j < n_boxes; j++) {
            box_lower_bounds[i * n_boxes + j] = j * box_width + x_min;
            box_upper_bounds[i * n_boxes + j] = (j + 1) * box_width + x_min;
            box_lower_bounds[n_total_boxes + i * n_boxes + j] = i * box_width + y_min;
            box_upper_bounds[n_total_boxes + i * n_boxes + j] = (i + 1) * box_width + y_min;
        }
    }
    double h = 1 / (double)n_interpolation_points;
    y_tilde_spacings[0] = h / 2;
    for (int i = 1; i < n_interpolation_points; i++) y_tilde_spacings[i] = y_tilde_spacings[i - 1] + h;

    int n1d = n_interpolation_points * n_boxes;
    int M = 2 * n1d;
    h = h * box_width;
    x_tilde[0] = x_min + h / 2;
    y_tilde[0] = y_min + h / 2;
    for (int i = 1; i < n1d; i++) {
        x_tilde[i] = x_tilde[i - 1] + h;
        y_tilde[i] = y_tilde[i - 1] + h;
    }
    double *kernel_tilde = (double *)calloc((size_t)M * M, sizeof(double));
    for (int i = 0; i < n1d; i++) {
        for (int j = 0; j < n1d; j++) {
            double tmp = orc_kernel_2d(kernel, y_tilde[0], x_tilde[0], y_tilde[i], x_tilde[j], df);
            kernel_tilde[(long)(n1d + i) * M + (n1d + j)] = tmp;
            kernel_tilde[(long)(n1d - i) * M + (n1d + j)] = tmp;
            kernel_tilde[(long)(n1d + i) * M + (n1d - j)] = tmp;
            kernel_tilde[(long)(n1d - i) * M + (n1d - j)] = tmp;
        }
    }
    orc_fft_r2c_2d(M, M, kernel_tilde, fft_kernel_tilde);
    free(kernel_tilde);
}

/* ---- nbodyfft.cpp:310-336 ---- */
void orc_interpolate(int n_interpolation_points, int N, const double *y_in_box,
                     const double *y_tilde_spacings, double *interpolated_values) {
    double *denominator = (double *)malloc(sizeof(double) * n_interpolation_points);
    for (int i = 0; i < n_interpolation_points; i++) {
        denominator[i] = 1;
        for (int j = 0; j < n_interpolation_points; j++)
            if (i != j) denominator[i] *= y_tilde_spacings[i] - y_tilde_spacings[j];
    }
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < n_interpolation_points; j++) {
            double v = 1;
            for (int k = 0; k < n_interpolation_points; k++)
                if (j != k) v *= y_in_box[i] - y_tilde_spacings[k];
            interpolated_values[(long)j * N + i] = v / denominator[j];
        }
    }
    free(denominator);
}

/* Step 3 of n_body_fft_2d runs under PARALLEL_FOR (nbodyfft.cpp:222-239). */
typedef struct {
    int N, n_terms, n_boxes, p;
    const int *point_box_idx;
    const double *xv, *yv, *y_tilde_values;
    double *potentialQij;
} step3_ctx;

static void step3_body(int lo, int hi, void *c_) {
    step3_ctx *c = (step3_ctx *)c_;
    int p = c->p, nb = c->n_boxes, N = c->N, nt = c->n_terms;
    for (int loop_i = lo; loop_i < hi; loop_i++) {
        int box_idx = c->point_box_idx[loop_i];
        int box_i = box_idx % nb;
        int box_j = box_idx / nb;
        for (int ii = 0; ii < p; ii++) {
            for (int jj = 0; jj < p; jj++) {
                for (int d = 0; d < nt; d++) {
                    long idx = (long)(box_i * p + ii) * (nb * p) + (box_j * p) + jj;
                    c->potentialQij[(long)loop_i * nt + d] += c->xv[(long)ii * N + loop_i] *
                                                              c->yv[(long)jj * N + loop_i] *
                                                              c->y_tilde_values[idx * nt + d];
                }
            }
        }
    }
}

/* ---- nbodyfft.cpp:73-248 ---- */
void orc_n_body_fft_2d(int N, int n_terms, const double *xs, const double *ys, const double *chargesQij,
                       int n_boxes, int n_interpolation_points, const double *box_lower_bounds,
                       const double *box_upper_bounds, const double *y_tilde_spacings,
                       const double *fft_kernel_tilde, double *potentialQij, unsigned nthreads) {
    int p = n_interpolation_points;
    int n_total_boxes = n_boxes * n_boxes;
    long total_interpolation_points = (long)n_total_boxes * p * p;
    double coord_min = box_lower_bounds[0];
    double box_width = box_upper_bounds[0] - box_lower_bounds[0];

    int *point_box_idx = (int *)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; i++) {
        int x_idx = (int)((xs[i] - coord_min) / box_width);
        int y_idx = (int)((ys[i] - coord_min) / box_width);
        if (x_idx >= n_boxes) x_idx = n_boxes - 1;
        else if (x_idx < 0) x_idx = 0;
        if (y_idx >= n_boxes) y_idx = n_boxes - 1;
        else if (y_idx < 0) y_idx = 0;
        point_box_idx[i] = y_idx * n_boxes + x_idx;
    }
    double *x_in_box = (double *)malloc(sizeof(double) * (size_t)N);
    double *y_in_box = (double *)malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; i++) {
        int box_idx = point_box_idx[i];
        double x_min = box_lower_bounds[box_idx];
        double y_min = box_lower_bounds[n_total_boxes + box_idx];
        x_in_box[i] = (xs[i] - x_min) / box_width;
        y_in_box[i] = (ys[i] - y_min) / box_width;
    }
    double *xv = (double *)malloc(sizeof(double) * (size_t)N * p);
    double *yv = (double *)malloc(sizeof(double) * (size_t)N * p);
    orc_interpolate(p, N, x_in_box, y_tilde_spacings, xv);
    orc_interpolate(p, N, y_in_box, y_tilde_spacings, yv);

    double *w = (double *)calloc((size_t)total_interpolation_points * n_terms, sizeof(double));
    for (int i = 0; i < N; i++) {
        int box_idx = point_box_idx[i];
        int box_j = box_idx / n_boxes;
        int box_i = box_idx % n_boxes;
        for (int ii = 0; ii < p; ii++) {
            for (int jj = 0; jj < p; jj++) {
                long idx = (long)(box_i * p + ii) * (n_boxes * p) + (box_j * p) + jj;
                for (int d = 0; d < n_terms; d++)
                    w[idx * n_terms + d] += yv[(long)jj * N + i] * xv[(long)ii * N + i] * chargesQij[(long)i * n_terms + d];
            }
        }
    }

    double *y_tilde_values = (double *)calloc((size_t)total_interpolation_points * n_terms, sizeof(double));
    int half = p * n_boxes;
    int M = 2 * half;
    long H = M / 2 + 1;
    double *mpol_sort = (double *)malloc(sizeof(double) * (size_t)total_interpolation_points);
    double *fft_input = (double *)calloc((size_t)M * M, sizeof(double));
    double *fft_w = (double *)malloc(sizeof(double) * 2 * (size_t)M * H);
    double *fft_output = (double *)calloc((size_t)M * M, sizeof(double));
    for (int d = 0; d < n_terms; d++) {
        for (long i = 0; i < total_interpolation_points; i++) mpol_sort[i] = w[i * n_terms + d];
        for (int i = 0; i < half; i++)
            for (int j = 0; j < half; j++) fft_input[(long)i * M + j] = mpol_sort[(long)i * half + j];
        orc_fft_r2c_2d(M, M, fft_input, fft_w);
        for (long i = 0; i < (long)M * H; i++) {
            double x_ = fft_w[2 * i], y_ = fft_w[2 * i + 1];
            double u_ = fft_kernel_tilde[2 * i], v_ = fft_kernel_tilde[2 * i + 1];
            fft_w[2 * i] = x_ * u_ - y_ * v_;
            fft_w[2 * i + 1] = x_ * v_ + y_ * u_;
        }
        orc_fft_c2r_2d(M, M, fft_w, fft_output);
        for (int i = 0; i < half; i++) {
            for (int j = 0; j < half; j++) {
                int row = half + i, col = half + j;
                mpol_sort[(long)i * half + j] = fft_output[(long)row * M + col] / (double)(M * M);
            }
        }
        for (long i = 0; i < (long)half * half; i++) y_tilde_values[i * n_terms + d] = mpol_sort[i];
    }
    free(fft_w);
    free(fft_input);
    free(fft_output);
    free(mpol_sort);

    step3_ctx c = {N, n_terms, n_boxes, p, point_box_idx, xv, yv, y_tilde_values, potentialQij};
    orc_parallel_for(nthreads, N, step3_body, &c);

    free(point_box_idx);
    free(xv);
    free(yv);
    free(w);
    free(y_tilde_values);
    free(x_in_box);
    free(y_in_box);
}

/* ---- nbodyfft.cpp:251-307 ---- */
void orc_precompute_1d(double y_min, double y_max, int n_boxes, int n_interpolation_points, int kernel,
                       double *box_lower_bounds, double *box_upper_bounds, double *y_tilde_spacing,
                       double *y_tilde, double *fft_kernel_vector, double df) {
    double box_width = (y_max - y_min) / (double)n_boxes;
    for (int b = 0; b < n_boxes; b++) {
        box_lower_bounds[b] = b * box_width + y_min;
        box_upper_bounds[b] = (b + 1) * box_width + y_min;
    }
    int total = n_interpolation_points * n_boxes;
    double h = 1 / (double)n_interpolation_points;
    y_tilde_spacing[0] = h / 2;
    for (int i = 1; i < n_interpolation_points; i++) y_tilde_spacing[i] = y_tilde_spacing[i - 1] + h;
    h = h * box_width;
    y_tilde[0] = y_min + h / 2;
    for (int i = 1; i < total; i++) y_tilde[i] = y_tilde[i - 1] + h;

    double *kv = (double *)calloc(2 * (size_t)2 * total, sizeof(double)); /* complex[2*total] */
    for (int i = 0; i < total; i++) kv[2 * (total + i)] = orc_kernel_1d(kernel, y_tilde[0], y_tilde[i], df);
    for (int i = 1; i < total; i++) kv[2 * i] = kv[2 * (2 * total - i)];
    orc_fft_c2c(2 * total, kv, fft_kernel_vector, 0);
    free(kv);
}

/* ---- nbodyfft.cpp:339-455 ---- */
void orc_n_body_fft_1d(int N, int n_terms, const double *Y, const double *chargesQij, int n_boxes,
                       int n_interpolation_points, const double *box_lower_bounds,
                       const double *box_upper_bounds, const double *y_tilde_spacings,
                       const double *y_tilde, const double *fft_kernel_vector, double *potentialsQij) {
    (void)y_tilde;
    int p = n_interpolation_points;
    int total = p * n_boxes;
    double coord_min = box_lower_bounds[0];
    double box_width = box_upper_bounds[0] - box_lower_bounds[0];
    int *point_box_idx = (int *)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; i++) {
        int box_idx = (int)((Y[i] - coord_min) / box_width);
        if (box_idx >= n_boxes) box_idx = n_boxes - 1;
        point_box_idx[i] = box_idx;
    }
    double *y_in_box = (double *)malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; i++) y_in_box[i] = (Y[i] - box_lower_bounds[point_box_idx[i]]) / box_width;
    double *iv = (double *)malloc(sizeof(double) * (size_t)p * N);
    orc_interpolate(p, N, y_in_box, y_tilde_spacings, iv);

    double *w = (double *)calloc((size_t)total * n_terms, sizeof(double));
    for (int i = 0; i < N; i++) {
        int bi = point_box_idx[i] * p;
        for (int k = 0; k < p; k++)
            for (int d = 0; d < n_terms; d++)
                w[(long)(bi + k) * n_terms + d] += iv[(long)k * N + i] * chargesQij[(long)i * n_terms + d];
    }
    double *emb = (double *)calloc((size_t)2 * total * n_terms, sizeof(double));
    for (int i = 0; i < total; i++)
        for (int d = 0; d < n_terms; d++) emb[(long)(total + i) * n_terms + d] = w[(long)i * n_terms + d];

    double *fw = (double *)malloc(sizeof(double) * 2 * (size_t)2 * total);
    double *ytv = (double *)calloc((size_t)total * n_terms, sizeof(double));
    for (int d = 0; d < n_terms; d++) {
        for (int i = 0; i < 2 * total; i++) {
            fw[2 * i] = emb[(long)i * n_terms + d];
            fw[2 * i + 1] = 0.0;
        }
        orc_fft_c2c(2 * total, fw, fw, 0);
        for (int i = 0; i < 2 * total; i++) {
            double x_ = fw[2 * i], y_ = fw[2 * i + 1];
            double u_ = fft_kernel_vector[2 * i], v_ = fft_kernel_vector[2 * i + 1];
            fw[2 * i] = x_ * u_ - y_ * v_;
            fw[2 * i + 1] = x_ * v_ + y_ * u_;
        }
        orc_fft_c2c(2 * total, fw, fw, 1);
        for (int i = 0; i < total; i++) ytv[(long)i * n_terms + d] = fw[2 * i] / (total * 2.0);
    }
    free(fw);
    for (int i = 0; i < N; i++) {
        int bi = point_box_idx[i] * p;
        for (int j = 0; j < p; j++)
            for (int d = 0; d < n_terms; d++)
                potentialsQij[(long)i * n_terms + d] += iv[(long)j * N + i] * ytv[(long)(bi + j) * n_terms + d];
    }
    free(point_box_idx);
    free(y_in_box);
    free(iv);
    free(w);
    free(ytv);
    free(emb);
}
