/*
 * ref_harness.cpp -- C entry points into the reference's OWN source (FIt-SNE 1.2.1: tsne.cpp,
 * nbodyfft.cpp, sptree.cpp, compiled from /root/reference/src by oracle/Makefile into
 * oracle/_ref/libfitsne_ref.so).  TEST INFRASTRUCTURE.
 *
 * Only reference functions that never reach FFTW are exposed: FFTW is absent from this image, so the
 * reference's FFT path (precompute_2d, n_body_fft_2d, precompute, nbodyfft, computeFftGradient*,
 * TSNE::run) is unbuildable here.  The library is linked with --gc-sections from objects built with
 * -fvisibility=hidden, so the FFT functions are discarded and the link is checked with --no-undefined:
 * nothing in it is a stand-in for a missing library.
 *
 * Private TSNE members are reached by redefining `private` around the include of tsne.h only
 * (standard headers are included first so the library's own declarations are untouched).
 */
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nbodyfft.h"
#define private public
#include "tsne.h"
#undef private

double squared_cauchy(double x, double y, double df);
double general_kernel(double x, double y, double df);
double squared_general_kernel(double x, double y, double df);
double squared_cauchy_2d(double x1, double x2, double y1, double y2, double df);
double general_kernel_2d(double x1, double x2, double y1, double y2, double df);
double squared_general_kernel_2d(double x1, double x2, double y1, double y2, double df);

#define API extern "C" __attribute__((visibility("default")))

/* nbodyfft.cpp:310-336 */
API void ref_interpolate(int p, int N, const double *y_in_box, const double *spacings, double *out) {
    interpolate(p, N, y_in_box, spacings, out);
}

/* tsne.cpp:69-94 */
API double ref_kernel_1d(int kind, double x, double y, double df) {
    if (kind == 0) return squared_cauchy(x, y, df);
    if (kind == 1) return general_kernel(x, y, df);
    return squared_general_kernel(x, y, df);
}
API double ref_kernel_2d(int kind, double x1, double x2, double y1, double y2, double df) {
    if (kind == 0) return squared_cauchy_2d(x1, x2, y1, y2, df);
    if (kind == 1) return general_kernel_2d(x1, x2, y1, y2, df);
    return squared_general_kernel_2d(x1, x2, y1, y2, df);
}

/* tsne.cpp:1394-1469 */
API double ref_distances2similarities(double *D, double *P, int N, int n, double perplexity, double sigma,
                                      int ifSquared) {
    TSNE t;
    return t.distances2similarities(D, P, N, n, perplexity, sigma, ifSquared != 0);
}

/* tsne.cpp:1474-1500 */
API double ref_distances2similarities_list(double *D, double *P, int N, int n, double perplexity, double sigma,
                                           int ifSquared, int list_len, double *list) {
    TSNE t;
    return t.distances2similarities(D, P, N, n, perplexity, sigma, ifSquared != 0, list_len, list);
}

/* tsne.cpp:1730-1828; returns nnz, outputs sized by the caller to 2 * row_P[N]. */
API int ref_symmetrize(const unsigned *row_P, const unsigned *col_P, const double *val_P, int N,
                       unsigned *sym_row, unsigned *sym_col, double *sym_val) {
    unsigned nnz = row_P[N];
    unsigned *r = (unsigned *)malloc(sizeof(unsigned) * (N + 1));
    unsigned *c = (unsigned *)malloc(sizeof(unsigned) * nnz);
    double *v = (double *)malloc(sizeof(double) * nnz);
    memcpy(r, row_P, sizeof(unsigned) * (N + 1));
    memcpy(c, col_P, sizeof(unsigned) * nnz);
    memcpy(v, val_P, sizeof(double) * nnz);
    TSNE t;
    t.symmetrizeMatrix(&r, &c, &v, N);
    int out = (int)r[N];
    memcpy(sym_row, r, sizeof(unsigned) * (N + 1));
    memcpy(sym_col, c, sizeof(unsigned) * out);
    memcpy(sym_val, v, sizeof(double) * out);
    free(r);
    free(c);
    free(v);
    return out;
}

/* tsne.cpp:1329-1355 with nthreads = 1 (no race) and current_sum_Q given. */
API double ref_evaluate_error_fft(unsigned *row_P, unsigned *col_P, double *val_P, double *Y, int N, int D,
                                  double sum_Q, double df) {
    TSNE t;
    t.current_sum_Q = sum_Q;
    return t.evaluateErrorFft(row_P, col_P, val_P, Y, N, D, 1, df);
}

/* tsne.cpp:1232-1282 */
API void ref_exact_gradient(double *P, double *Y, int N, int D, double *dC, double df) {
    TSNE t;
    t.computeExactGradient(P, Y, N, D, dC, df);
}

/* tsne.cpp:1286-1326 */
API double ref_evaluate_error_exact(double *P, double *Y, int N, int D, double df) {
    TSNE t;
    return t.evaluateError(P, Y, N, D, df);
}

/* tsne.cpp:1832-1848 */
API void ref_squared_distances(double *X, int N, int D, double *DD) {
    TSNE t;
    t.computeSquaredEuclideanDistance(X, N, D, DD);
}

/* tsne.cpp:1503-1530 (dense overload used by the exact mode); P[N*N] */
API void ref_gaussian_perplexity_dense(double *X, int N, int D, double *P, double perplexity, double sigma,
                                       int list_len, double *list) {
    TSNE t;
    t.computeGaussianPerplexity(X, N, D, P, perplexity, sigma, list_len, list);
}

/* tsne.cpp:596-640 (Barnes-Hut gradient over the reference's SPTree), nthreads = 1 */
API void ref_bh_gradient(unsigned *row_P, unsigned *col_P, double *val_P, double *Y, int N, int D, double theta,
                         double *dC) {
    TSNE t;
    t.computeGradient(nullptr, row_P, col_P, val_P, Y, N, D, dC, theta, 1);
}

/* tsne.cpp:1359-1390 (Barnes-Hut KL), nthreads = 1 */
API double ref_bh_error(unsigned *row_P, unsigned *col_P, double *val_P, double *Y, int N, int D, double theta) {
    TSNE t;
    return t.evaluateError(row_P, col_P, val_P, Y, N, D, theta, 1);
}

/* tsne.cpp:1851-1876 */
API void ref_zero_mean(double *X, int N, int D) {
    TSNE t;
    t.zeroMean(X, N, D);
}

/* tsne.cpp:1643-1726 (VP tree kNN + perplexity rows, before symmetrisation).
 * row_P[N+1], col_P[N*K], val_P[N*K] are filled from the reference's own allocation. */
API void ref_gaussian_perplexity_vptree(double *X, int N, int D, double perplexity, int K, double sigma,
                                        unsigned nthreads, int list_len, double *list, unsigned seed,
                                        unsigned *row_P, unsigned *col_P, double *val_P) {
    srand(seed);
    TSNE t;
    unsigned *r = nullptr, *c = nullptr;
    double *v = nullptr;
    t.computeGaussianPerplexity(X, N, D, &r, &c, &v, perplexity, K, sigma, nthreads, list_len, list);
    memcpy(row_P, r, sizeof(unsigned) * (N + 1));
    memcpy(col_P, c, sizeof(unsigned) * (size_t)N * K);
    memcpy(val_P, v, sizeof(double) * (size_t)N * K);
    free(r);
    free(c);
    free(v);
}

/* tsne.cpp:1895-2020: parse a data.dat; scalars are returned in two arrays in file order.
 * ints: n, d, no_dims, max_iter, stop_lying_iter, mom_switch_iter, K, nbody_algo, knn_algo,
 *       no_momentum_during_exag, n_trees, search_k, start_late_exag_iter, nterms, min_num_intervals,
 *       rand_seed, load_affinities, skip_random_init, perplexity_list_length
 * dbls: theta, perplexity, momentum, final_momentum, learning_rate, max_step_norm, sigma,
 *       early_exag_coeff, late_exag_coeff, intervals_per_integer, df
 * X (n*d) and Y (n*no_dims) are copied out when the caller's buffers are non-null. */
API int ref_load_data(const char *path, int *ints, double *dbls, double *X, double *Y, double *plist) {
    TSNE t;
    double *data = nullptr, *Yp = nullptr, *pl = nullptr;
    int n = 0, d = 0, no_dims = 0, max_iter = 0, stop = 0, mom = 0, K = 0, nbody = 0, knn = 0, nomom = 0;
    int n_trees = 0, search_k = 0, late_iter = 0, nterms = 0, min_int = 0, seed = 0, load_aff = 0, pll = 0;
    double theta = 0, perp = 0, momentum = 0, fmom = 0, lr = 0, msn = 0, sigma = 0, early = 0, late = 0, ipi = 0,
           df = 1.0;
    bool skip = false;
    bool ok = t.load_data(path, &data, &Yp, &n, &d, &no_dims, &theta, &perp, &seed, &max_iter, &stop, &mom,
                          &momentum, &fmom, &lr, &K, &sigma, &nbody, &knn, &early, &nomom, &n_trees, &search_k,
                          &late_iter, &late, &nterms, &ipi, &min_int, &skip, &load_aff, &pll, &pl, &df, &msn);
    if (!ok) return -1;
    int iv[19] = {n, d, no_dims, max_iter, stop, mom, K, nbody, knn, nomom, n_trees, search_k, late_iter,
                  nterms, min_int, seed, load_aff, skip ? 1 : 0, perp == 0 ? pll : 0};
    double dv[11] = {theta, perp, momentum, fmom, lr, msn, sigma, early, late, ipi, df};
    memcpy(ints, iv, sizeof(iv));
    memcpy(dbls, dv, sizeof(dv));
    if (X) memcpy(X, data, sizeof(double) * (size_t)n * d);
    if (Y && skip) memcpy(Y, Yp, sizeof(double) * (size_t)n * no_dims);
    if (plist && perp == 0 && pl) memcpy(plist, pl, sizeof(double) * pll);
    free(data);
    free(Yp);
    if (perp == 0) free(pl);
    return 0;
}

/* tsne.cpp:2024-2038 */
API void ref_save_data(const char *path, double *Y, double *costs, int n, int d, int max_iter) {
    TSNE t;
    t.save_data(path, Y, costs, n, d, max_iter);
}
