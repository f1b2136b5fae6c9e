/*
 * fitsne_oracle.h -- FP64 CPU restatement of FIt-SNE's per-iteration gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU product path is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product (fit-sne_amd/) does not link it.
 *
 * Every function restates one reference function (FIt-SNE 1.2.1, file:line given beside it) with the
 * same loop structure, argument order and numerical quirks.  FFTW (absent from this image) is replaced
 * by the oracle's own mixed-radix FFT (orc_fft_*), which follows FFTW's conventions: unnormalised,
 * exp(-2 pi i jk/n) forward, r2c output n0 x (n1/2+1), c2r ignoring the imaginary parts that a
 * Hermitian input leaves on the self-conjugate bins.
 *
 * Parity pinning (what checks each function; details in tests/test_oracle_*.py):
 *   orc_fft_*                 DFT definition (direct O(n^2) sum), inverse round trip, Parseval
 *   orc_interpolate           reference nbodyfft.cpp:310 built from source (oracle/_ref), Lagrange invariants
 *   orc_precompute_2d / _1d   closed forms for nodes/bounds; kernel spectrum vs direct DFT of the embedding
 *   orc_n_body_fft_2d / _1d   circulant step == direct linear convolution on the node lattice; end to end
 *                             within interpolation error of the exact O(N^2) sum
 *   orc_compute_fft_gradient  repulsion within interpolation error of reference computeExactGradient
 *                             (oracle/_ref); attraction == brute force; Z identity
 *   orc_evaluate_error_fft    reference evaluateErrorFft (oracle/_ref, nthreads=1)
 *   orc_distances2similarities reference distances2similarities (oracle/_ref), bit-exact
 *   orc_symmetrize            reference symmetrizeMatrix (oracle/_ref), bit-exact
 *   orc_gaussian_perplexity_exact_knn  reference VP-tree computeGaussianPerplexity (oracle/_ref)
 *   orc_zero_mean             reference zeroMean (oracle/_ref)
 *   orc_squared_distances     reference computeSquaredEuclideanDistance (oracle/_ref), bit-exact
 *   orc_gaussian_perplexity_dense  reference dense computeGaussianPerplexity (oracle/_ref), bit-exact
 *   orc_exact_symmetrize_normalize  symmetry, unit total, brute-force (P + P^T) / sum on small inputs
 *   orc_exact_gradient        reference computeExactGradient (oracle/_ref), bit-exact; finite differences
 *                             of the exact KL
 *   orc_evaluate_error_exact  reference evaluateError (oracle/_ref), bit-exact
 *   orc_bh_gradient / orc_bh_error  reference computeGradient (SPTree) / evaluateError (oracle/_ref),
 *                             bit-exact; the exact O(N^2) gradient as theta -> 0
 *   orc_update                elementwise closed form of tsne.cpp:478-514
 *   orc_run_fft / _bh / _exact one loop (exaggeration switches, momentum switch, no_momentum_during_exag,
 *                             step clip, KL correction) around the three gradient sources; the Barnes-Hut
 *                             and exact forms are compared with the reference's own TSNE::run
 *                             (oracle/_ref/libfitsne_ref_run.so, tests/test_oracle_run.py), which pins the
 *                             loop the FFT form shares
 */
#ifndef FITSNE_ORACLE_H
#define FITSNE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel ids standing in for the reference's function pointers (tsne.cpp:69-94, nbodyfft.h:13-15). */
enum {
    ORC_SQUARED_CAUCHY = 0,         /* (1 + r^2)^-2            squared_cauchy[_2d]          */
    ORC_GENERAL = 1,                /* (1 + r^2/df)^-df        general_kernel[_2d]          */
    ORC_SQUARED_GENERAL = 2         /* (1 + r^2/df)^-(df+1)    squared_general_kernel[_2d]  */
};

double orc_kernel_1d(int kind, double x, double y, double df);
double orc_kernel_2d(int kind, double x1, double x2, double y1, double y2, double df);

/* ---- FFT (stand-in for the FFTW calls at nbodyfft.cpp:64-68, :164-168, :301-304, :400-404) ---- */
/* complex arrays are interleaved (re, im) doubles, layout-compatible with std::complex<double>. */
void orc_fft_c2c(int n, const double *in, double *out, int inverse);
void orc_fft_r2c_2d(int n0, int n1, const double *in, double *out);
void orc_fft_c2r_2d(int n0, int n1, const double *in, double *out);
void orc_dft_c2c_naive(int n, const double *in, double *out, int inverse); /* pinning reference */

/* ---- PARALLEL_FOR (parallel_for.h:26-41): static chunks [t*n/T, (t+1)*n/T) ---- */
typedef void (*orc_body_fn)(int lo, int hi, void *ctx);
void orc_parallel_for(unsigned nthreads, int n, orc_body_fn fn, void *ctx);

/* ---- nbodyfft.cpp ---- */
void orc_precompute_2d(double x_max, double x_min, double y_max, double y_min, int n_boxes,
                       int n_interpolation_points, int kernel, double *box_lower_bounds,
                       double *box_upper_bounds, double *y_tilde_spacings, double *y_tilde,
                       double *x_tilde, double *fft_kernel_tilde, double df);
void orc_n_body_fft_2d(int N, int n_terms, const double *xs, const double *ys, const double *chargesQij,
                       int n_boxes, int n_interpolation_points, const double *box_lower_bounds,
                       const double *box_upper_bounds, const double *y_tilde_spacings,
                       const double *fft_kernel_tilde, double *potentialQij, unsigned nthreads);
void orc_precompute_1d(double y_min, double y_max, int n_boxes, int n_interpolation_points, int kernel,
                       double *box_lower_bounds, double *box_upper_bounds, double *y_tilde_spacing,
                       double *y_tilde, double *fft_kernel_vector, double df);
void orc_n_body_fft_1d(int N, int n_terms, const double *Y, const double *chargesQij, int n_boxes,
                       int n_interpolation_points, const double *box_lower_bounds,
                       const double *box_upper_bounds, const double *y_tilde_spacings,
                       const double *y_tilde, const double *fft_kernel_vector, double *potentialsQij);
void orc_interpolate(int n_interpolation_points, int N, const double *y_in_box,
                     const double *y_tilde_spacings, double *interpolated_values);

/* ---- grid geometry helpers (tsne.cpp:1042-1077 and :656-663) ---- */
/* 2D min/max with the reference's `else if` rule (tsne.cpp:1045-1048). */
void orc_minmax_2d_quirk(const double *Y, int N, double *min_coord, double *max_coord);
/* n_boxes_per_dim with truncation and the 2D rounding table (tsne.cpp:1065-1077). */
int orc_n_boxes_2d(double min_coord, double max_coord, double intervals_per_integer, int min_num_intervals);
int orc_n_boxes_1d(double y_min, double y_max, double intervals_per_integer, int min_num_intervals);

/* ---- gradients (tsne.cpp:645-1169) ---- */
/* Dispatch exactly like tsne.cpp:444-464: D in {1,2} x (df == 1 or not).  dC[N*D] is overwritten,
 * *sum_Q receives current_sum_Q.  Returns 0, or -1 for D outside {1,2}. */
int orc_compute_fft_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                             const double *Y, int N, int D, double *dC, int n_interpolation_points,
                             double intervals_per_integer, int min_num_intervals, unsigned nthreads,
                             double df, double *sum_Q);
/* The four variants, callable directly. */
void orc_grad_2d(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                 double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double *sum_Q);
void orc_grad_2d_df(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double df,
                    double *sum_Q);
void orc_grad_1d(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                 double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double *sum_Q);
void orc_grad_1d_df(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    double *dC, int p, double ipi, int min_num_intervals, unsigned nthreads, double df,
                    double *sum_Q);
/* Attractive term alone (tsne.cpp:1121-1137, :965-980, :826-837, :721-732), pos_f[N*D]. */
void orc_attractive(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y,
                    int N, int D, double df, unsigned nthreads, double *pos_f);
/* Exact O(N^2) repulsion sum_j q^((df+1)/df) (y_i - y_j) / Z with Z over i != j (tsne.cpp:1172-1228)
 * and the exact normaliser; used only to bound interpolation error in tests. */
void orc_exact_repulsion(const double *Y, int N, int D, double df, double *neg_f, double *sum_Q);

/* ---- optimiser and error (tsne.cpp:478-544, :1329-1355, :1851-1876) ---- */
void orc_zero_mean(double *X, int N, int D);
/* One update step exactly as tsne.cpp:478-514 (gains, momentum, clip, Y += uY or Y -= dY). */
void orc_update(double *Y, double *uY, double *gains, const double *dY, int N, int D, double momentum,
                double learning_rate, double max_step_norm, int plain_descent);
/* Race-free evaluateErrorFft: per-row partials summed in row order (tsne.cpp:1329-1355). */
double orc_evaluate_error_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                              const double *Y, int N, int D, double sum_Q, double df);

typedef struct {
    int max_iter, stop_lying_iter, mom_switch_iter, start_late_exag_iter, no_momentum_during_exag;
    int n_interpolation_points, min_num_intervals;
    double momentum, final_momentum, learning_rate, early_exag_coeff, late_exag_coeff;
    double max_step_norm, df, intervals_per_integer;
    unsigned nthreads;
} orc_run_params;
/* The FFT branch of TSNE::run's hot loop (tsne.cpp:389-577) on a given symmetric normalised P
 * (val_P is copied; exaggeration is applied to the copy in place like the reference).  Y is the
 * initial embedding and receives the result; costs[max_iter] gets the KL entries. */
int orc_run_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N, int no_dims,
                double *Y, const orc_run_params *prm, double *costs);
/* The same loop with the Barnes-Hut gradient and KL (nbody_algorithm = 1, tsne.cpp:458-460, :556-558), and
 * with the exact gradient over a dense P[N*N] (theta = 0, tsne.cpp:442-443, :552-553; early_exag_coeff
 * must be given: its automatic value needs the CSR).  Returns 0, or -1 for an unsupported request. */
int orc_run_bh(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N, int no_dims,
               double *Y, const orc_run_params *prm, double theta, double *costs);
int orc_run_exact(const double *P, int N, int no_dims, double *Y, const orc_run_params *prm, double *costs);

/* ---- affinities (tsne.cpp:1394-1500, :1643-1726, :1730-1828) ---- */
double orc_distances2similarities(const double *D, double *P, int N, int n, double perplexity,
                                  double sigma, int ifSquared);
/* Same bisection with a caller-chosen starting beta; *iters receives the number of entropy
 * evaluations (the reference's loop counter).  beta0 = 1 reproduces the reference exactly. */
double orc_d2s_bisect(const double *D, double *P, int N, double perplexity, double beta0, int ifSquared,
                      int *iters, double *H_out);
double orc_distances2similarities_list(const double *D, double *P, int N, int n, double perplexity,
                                       double sigma, int ifSquared, int perplexity_list_length,
                                       const double *perplexity_list);
/* Exact kNN (brute force, Euclidean as vptree.h computes it), excluding self; ascending.
 * idx[N*K], dist[N*K] (non-squared). Ties keep the lower index first. */
void orc_knn_exact(const double *X, int N, int D, int K, int32_t *idx, double *dist, unsigned nthreads);
/* computeGaussianPerplexity's per-row P from kNN output (tsne.cpp:1684-1706). */
void orc_rows_from_knn(const int32_t *idx, const double *dist, int N, int K, double perplexity,
                       double sigma, int perplexity_list_length, const double *perplexity_list,
                       uint32_t *row_P, uint32_t *col_P, double *val_P);
/* symmetrizeMatrix (tsne.cpp:1730-1828).  Returns nnz; outputs must hold 2*row_P[N] entries. */
int orc_symmetrize(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N,
                   uint32_t *sym_row_P, uint32_t *sym_col_P, double *sym_val_P);

/* ---- exact t-SNE, dense P[N*N] row-major (tsne.cpp:171-229, :1232-1326, :1503-1530, :1832-1848) ---- */
void orc_squared_distances(const double *X, int N, int D, double *DD);
void orc_gaussian_perplexity_dense(const double *X, int N, int D, double perplexity, double sigma,
                                   int perplexity_list_length, const double *perplexity_list, double *P,
                                   unsigned nthreads);
double orc_exact_symmetrize_normalize(double *P, int N);
void orc_exact_gradient(const double *P, const double *Y, int N, int D, double df, double *dC, double *sum_Q);
double orc_evaluate_error_exact(const double *P, const double *Y, int N, int D, double df);

/* ---- Barnes-Hut (nbody_algo = 1; sptree.cpp, tsne.cpp:596-640, :1359-1390), D <= 3 ---- */
void orc_bh_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                     int D, double theta, double *dC, double *sum_Q);
double orc_bh_error(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, const double *Y, int N,
                    int D, double theta);

#ifdef __cplusplus
}
#endif
#endif
