/*
 * fitsne_abi.h -- C-ABI of the MI355X-native FIt-SNE gradient library (libfitsne.so).
 *
 * Plain pointers and sizes only.  Every call returns an int status (FITSNE_OK or a negative code)
 * where the reference returns void; out-parameters carry what the reference returned or stored in
 * TSNE members.  Two tiers:
 *
 *  1. Mirror tier -- host pointers, FP64 in and out, synchronous, same parameter order and meaning
 *     as the reference functions named beside each entry point (FIt-SNE 1.2.1).  Runs the FP64
 *     "plain" GPU path, whose stages mirror the reference's.
 *  2. Resident tier -- a context owning device-resident P, Y, momentum and gains, so that one
 *     optimisation iteration moves no data over PCIe.  Offers the FP64 plain path and the FP32
 *     form (FP32 spread/FFT/interpolation, FP64 normalisation sum Z and grid geometry).
 *
 * The library is not re-entrant on one context; distinct contexts may be used from distinct
 * host threads.  No host array is retained after a call returns.
 */
#ifndef FITSNE_ABI_H
#define FITSNE_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FITSNE_OK 0
#define FITSNE_E_ARG (-1)         /* invalid argument (sizes, dims, null pointers)          */
#define FITSNE_E_ALLOC (-2)       /* device or host allocation failed                        */
#define FITSNE_E_HIP (-3)         /* a HIP runtime call failed (message: fitsne_last_error)  */
#define FITSNE_E_UNSUPPORTED (-4) /* valid for the reference but outside this build's limits */
#define FITSNE_E_NODEVICE (-5)    /* no usable GPU                                            */

/* kernel ids for the precompute calls: the reference passes function pointers
 * (nbodyfft.h:13-15) to squared_cauchy[_2d], general_kernel[_2d], squared_general_kernel[_2d]
 * (tsne.cpp:69-94). */
#define FITSNE_KERNEL_SQUARED_CAUCHY 0
#define FITSNE_KERNEL_GENERAL 1
#define FITSNE_KERNEL_SQUARED_GENERAL 2

/* precisions of the resident tier */
#define FITSNE_FP64_PLAIN 0 /* FP64 throughout, stages as in the reference               */
#define FITSNE_FP32 1       /* FP32 grid/FFT/interpolation, FP64 Z/geometry (product form) */
#define FITSNE_FP64_EXACT 2 /* exact t-SNE (theta = 0): dense P, O(N^2) gradient and KL, FP64 */
#define FITSNE_FP64_BH 3    /* Barnes-Hut (nbody_algo = 1): the reference's SPTree, FP64       */

/* Text of the last error on this thread (never NULL). */
const char *fitsne_last_error(void);
/* Number of visible GPUs (0 when none); no device is initialised beyond enumeration. */
int fitsne_device_count(void);
/* Library version string ("fitsne-mi355x <semver> (reference 1.2.1)"). */
const char *fitsne_version(void);

/* ====================================================================================== */
/* Mirror tier                                                                             */
/* ====================================================================================== */

/* Replaces TSNE::computeFftGradient / computeFftGradientVariableDf / computeFftGradientOneD /
 * computeFftGradientOneDVariableDf (tsne.cpp:1027, :860, :758, :647; declared tsne.h:71-94), with the
 * dispatch of TSNE::run (tsne.cpp:444-464): D in {1,2}, df == 1 or not.  The unused dense P argument
 * of the reference is dropped.  dC[N*D] is overwritten; *sum_Q receives what the reference stores
 * in TSNE::current_sum_Q (tsne.h:63).  nthreads is accepted for signature parity and ignored. */
int fitsne_compute_fft_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                                const double *Y, int N, int D, double *dC, int n_interpolation_points,
                                double intervals_per_integer, int min_num_intervals, unsigned nthreads,
                                double df, double *sum_Q);
/* Same with 64-bit row offsets and 32-bit columns (for nnz >= 2^32 inputs). */
int fitsne_compute_fft_gradient_i64(const int64_t *row_P, const int32_t *col_P, const double *val_P,
                                    const double *Y, int N, int D, double *dC, int n_interpolation_points,
                                    double intervals_per_integer, int min_num_intervals, unsigned nthreads,
                                    double df, double *sum_Q);

/* Replaces precompute_2d (nbodyfft.cpp:9-70, nbodyfft.h:17-19).  Box bounds, node spacings and node
 * coordinates are written exactly as the reference lays them out.  fft_kernel_tilde (caller-sized
 * for the reference: 2n x (n+1) complex doubles, n = n_boxes * n_interpolation_points) receives this
 * library's kernel spectrum, which n_body_fft_2d consumes: like the reference, callers treat it as
 * opaque between the two calls. */
int fitsne_precompute_2d(double x_max, double x_min, double y_max, double y_min, int n_boxes,
                         int n_interpolation_points, int kernel, double *box_lower_bounds,
                         double *box_upper_bounds, double *y_tilde_spacings, double *y_tilde,
                         double *x_tilde, double *fft_kernel_tilde, double df);

/* Replaces n_body_fft_2d (nbodyfft.cpp:73-248, nbodyfft.h:21-23).  potentialQij[N*n_terms] is
 * ACCUMULATED into (+=) as in the reference; n_terms in [1, 4]. */
int fitsne_n_body_fft_2d(int N, int n_terms, const double *xs, const double *ys, const double *chargesQij,
                         int n_boxes, int n_interpolation_points, const double *box_lower_bounds,
                         const double *box_upper_bounds, const double *y_tilde_spacings,
                         const double *fft_kernel_tilde, double *potentialQij, unsigned nthreads);

/* Replaces precompute (nbodyfft.cpp:251-307, nbodyfft.h:25-27); fft_kernel_vector (2n complex
 * doubles in the reference) receives this library's spectrum, consumed by n_body_fft_1d. */
int fitsne_precompute_1d(double y_min, double y_max, int n_boxes, int n_interpolation_points, int kernel,
                         double *box_lower_bounds, double *box_upper_bounds, double *y_tilde_spacing,
                         double *y_tilde, double *fft_kernel_vector, double df);

/* Replaces nbodyfft (nbodyfft.cpp:339-455, nbodyfft.h:29-31); potentialsQij accumulated (+=). */
int fitsne_n_body_fft_1d(int N, int n_terms, const double *Y, const double *chargesQij, int n_boxes,
                         int n_interpolation_points, const double *box_lower_bounds,
                         const double *box_upper_bounds, const double *y_tilde_spacings,
                         const double *y_tilde, const double *fft_kernel_vector, double *potentialsQij);

/* Replaces interpolate (nbodyfft.cpp:310-336, nbodyfft.h:33-34): out[p*N], SoA (j*N + i). */
int fitsne_interpolate(int n_interpolation_points, int N, const double *y_in_box, const double *y_tilde_spacings,
                       double *interpolated_values);

/* Replaces TSNE::evaluateErrorFft (tsne.cpp:1329-1355; tsne.h:102) given current_sum_Q, without the
 * reference's unsynchronised `C += temp` (tsne.cpp:1349): deterministic FP64 reduction. */
int fitsne_evaluate_error_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                              const double *Y, int N, int D, double sum_Q, double df, double *C);

/* No counterpart in the reference: fitsne_evaluate_error_fft for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, result and errors, bitwise the same C.  It allocates nothing: the staged CSR, Y and the
 * partial sums live in a fixed 32 MiB device array per GPU, held until the process ends; sizes whose
 * staging needs more (about 4 (N+1) + 12 nnz + 8 N D + N bytes) return FITSNE_E_ARG.  Runs on the
 * calling thread's current device; calls within one process are serialised. */
int fitsne_evaluate_error_fft_noalloc(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                                      const double *Y, int N, int D, double sum_Q, double df, double *C);

/* Replaces TSNE::distances2similarities (tsne.cpp:1394-1469; tsne.h:105) for a batch of points:
 * D[n_points*K] distances (row i = point i), P[n_points*K] row-normalised affinities, beta[n_points].
 * beta0 (nullable) gives per-point starting betas for the bisection (the reference uses 1.0);
 * iters (nullable) receives the reference's loop counter per point.  perplexity <= 0 selects the
 * fixed-sigma branch.  The FP64 bisection is the reference's, step for step. */
int fitsne_distances2similarities_batch(const double *D, double *P, int n_points, int K, double perplexity,
                                        double sigma, int ifSquared, const double *beta0, double *beta,
                                        int *iters);

/* Exact K nearest neighbours of every row of X[N*D] (row-major) among the other rows, for the
 * affinity build of TSNE::run: replaces the VP-tree search of computeGaussianPerplexity
 * (tsne.cpp:1643-1726; Annoy's approximate search, tsne.cpp:1531-1640, is served by the same exact
 * search).  idx[N*K] and dist[N*K] per row ascending by (distance, index); the distance is
 * sqrt(sum_d (x_d - y_d)^2) in order d = 0..D-1 without fused multiply-adds (vptree.h:97-107).
 * K <= min(N - 1, 1000); X must be finite. */
int fitsne_knn_exact(const double *X, int N, int D, int K, int32_t *idx, double *dist);

/* The same neighbours as fitsne_knn_exact, bit for bit, for large N: candidates from FP32 MFMA distances
 * (a sampled threshold, then every point below it), their FP64 distances in the reference's order, and
 * the FP64 kernel for any query whose candidates cannot be shown to contain its K nearest (their count,
 * or the K-th distance against the threshold minus the FP32 error bound).  D <= 128, K <= 256;
 * recomputed (nullable) receives the number of such queries.  Same role as fitsne_knn_exact: the exact
 * kNN of TSNE::run (tsne.cpp:1643-1726, vptree.h). */
int fitsne_knn_mfma(const double *X, int N, int D, int K, int32_t *idx, double *dist, long long *recomputed);

/* symmetrizeMatrix (tsne.cpp:1730-1828) on the GPU: the same output rows, entry order and values
 * ((P[n][c] + P[c][n]) / 2, or P[n][c] / 2 for an entry without partner).  Input: a CSR with N rows,
 * row_P[N] < 2^30 entries, no repeated (row, col) entry.  sym_col_P / sym_val_P must hold 2 * row_P[N]
 * entries; *sym_nnz (nullable) receives the output size sym_row_P[N].  The normalisation by the total
 * (tsne.cpp:327-330) is the caller's, as in the reference. */
int fitsne_symmetrize(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N,
                      uint32_t *sym_row_P, uint32_t *sym_col_P, double *sym_val_P, long long *sym_nnz);

/* ====================================================================================== */
/* Learned warm start of the perplexity calibration (no reference counterpart: the reference */
/* starts every bisection of distances2similarities, tsne.cpp:1394-1469, at beta = 1).      */
/* ====================================================================================== */

/* A small MLP (BF16 MFMA, FP32 accumulation) mapping a point's K <= 96 ascending kNN distances to a
 * starting beta for fitsne_distances2similarities_batch (beta0), whose unchanged bisection remains the
 * judge of the entropy error.  Features: standardised log(d^2); layers 96 -> 64 -> 64 -> 1 (ReLU);
 * output log(beta).  Parameter vector (FP32, 10,625 values): mu[96] sd[96] W1[96][64] b1[64] W2[64][64]
 * b2[64] w3[64] b3. */
typedef struct fitsne_beta_model fitsne_beta_model;
int fitsne_beta_model_create(fitsne_beta_model **model, int K, unsigned long long seed);
int fitsne_beta_model_destroy(fitsne_beta_model *model);
/* Train on n points (dist[n*K] as fitsne_knn_exact returns them, beta[n] from the bisection): feature
 * scales from these data, then `epochs` passes of Adam (learning rate lr) over shuffled mini-batches of
 * `batch` points, mean squared error on log(beta).  *final_loss (nullable): loss over all n after. */
int fitsne_beta_model_fit(fitsne_beta_model *model, const double *dist, const double *beta, int n, int epochs,
                          int batch, double lr, unsigned long long seed, double *final_loss);
/* beta0[n] = exp(model output) for dist[n*K]. */
int fitsne_beta_model_predict(fitsne_beta_model *model, const double *dist, int n, double *beta0);
/* Mean squared error over the n points and its gradient with respect to W1 .. b3 (grad[10,433], the
 * parameter layout above without mu and sd), no update. */
int fitsne_beta_model_gradient(fitsne_beta_model *model, const double *dist, const double *beta, int n,
                               float *grad, double *loss);
int fitsne_beta_model_params(fitsne_beta_model *model, float *params);
int fitsne_beta_model_set_params(fitsne_beta_model *model, const float *params);

/* ====================================================================================== */
/* Resident tier                                                                           */
/* ====================================================================================== */

typedef struct fitsne_ctx fitsne_ctx;

/* Optimiser parameters of one iteration (tsne.cpp:478-514). */
typedef struct {
    double momentum;          /* momentum of this iteration                                      */
    double learning_rate;     /* learning_rate                                                    */
    double max_step_norm;     /* clip per-point step norm when > 0 (tsne.cpp:498-511)             */
    int plain_descent;        /* 1: Y -= dY without gains/momentum (no_momentum_during_exag phase) */
    int reserved;
} fitsne_step_params;

/* Create a context on `device` for an embedding of dimension no_dims (1 or 2; 1..4 for
 * FITSNE_FP64_EXACT, 1..3 for FITSNE_FP64_BH) with the FFT settings of the run (n_interpolation_points = the wrapper's nterms,
 * intervals_per_integer, min_num_intervals; unused by the exact path) and the kernel's degree of
 * freedom df. */
int fitsne_ctx_create(fitsne_ctx **ctx, int device, int precision, int no_dims, int n_interpolation_points,
                      double intervals_per_integer, int min_num_intervals, double df);
int fitsne_ctx_destroy(fitsne_ctx *ctx);

/* Upload the symmetric affinities in CSR form (tsne.cpp:327-330 output).  N may differ from a
 * previous call.  The FP32 context stages val_P in FP32 once. */
int fitsne_set_P(fitsne_ctx *ctx, int N, const uint32_t *row_P, const uint32_t *col_P, const double *val_P);
int fitsne_set_P_i64(fitsne_ctx *ctx, int N, const int64_t *row_P, const int32_t *col_P, const double *val_P);
/* Exact t-SNE (theta = 0), FITSNE_FP64_EXACT contexts only; no_dims 1..4.  P is the dense N x N
 * matrix of the reference's exact mode (tsne.cpp:171-229), kept on the device.
 * fitsne_exact_affinities replaces computeGaussianPerplexity's dense overload (tsne.cpp:1503-1530:
 * squared distances of tsne.cpp:1832-1848, each row calibrated with its own entry set to DBL_MIN;
 * perplexity > 0: that perplexity; perplexity == 0: the average over perplexity_list; perplexity < 0:
 * the fixed width sigma) followed by the symmetrisation P[n][m] = P[m][n] = P[n][m] + P[m][n] and the
 * division by the total (tsne.cpp:196-216).  X[N*D] is the caller's (already centred and scaled)
 * data.  P_out (nullable) receives the N*N matrix (the P.dat contents); sum_P (nullable) the total
 * before normalisation.  fitsne_set_P_dense uploads a dense P (the P.dat route, tsne.cpp:172-189).
 * On such a context fitsne_gradient / fitsne_step compute computeExactGradient (tsne.cpp:1232-1282;
 * sum_Q as there) and fitsne_kl computes evaluateError (tsne.cpp:1286-1326) at the resident Y. */
int fitsne_exact_affinities(fitsne_ctx *ctx, const double *X, int N, int D, double perplexity, double sigma,
                            int perplexity_list_length, const double *perplexity_list, double *P_out,
                            double *sum_P);
int fitsne_set_P_dense(fitsne_ctx *ctx, int N, const double *P);

/* Barnes-Hut (nbody_algo = 1, theta > 0), FITSNE_FP64_BH contexts only; no_dims 1..3; CSR P as
 * fitsne_set_P.  fitsne_gradient / fitsne_step compute computeGradient (tsne.cpp:596-640): the SPTree of
 * sptree.cpp built each iteration on the host in the reference's insertion order, its walk
 * (computeNonEdgeForces) and the edge forces on the GPU, bitwise the reference's result.  fitsne_kl
 * computes the Barnes-Hut evaluateError (tsne.cpp:1359-1390) at the current Y.  The kernel is the
 * Cauchy one whatever df (the reference's Barnes-Hut path has no df).  Default theta 0.5. */
int fitsne_set_theta(fitsne_ctx *ctx, double theta);

/* Scalar multiplier on P (early/late exaggeration, tsne.cpp:389-412, :534-543).  The FP64 contexts
 * (plain, exact, Barnes-Hut) rescale their copy of P in place like the reference; the FP32 context keeps
 * a scalar. */
int fitsne_set_exaggeration(fitsne_ctx *ctx, double factor);

/* Spatial order (FP32 context): points are kept internally in the Morton order of their cell on a
 * 512x512 lattice over the embedding, so that neighbour gathers and the spread stay in cache.  All host
 * arrays remain in the caller's order.  fitsne_set_Y reorders when enabled (default) and N >= 65536;
 * fitsne_reorder re-sorts now (e.g. once clusters have formed).  No reference counterpart. */
int fitsne_set_reorder(fitsne_ctx *ctx, int enable);
int fitsne_reorder(fitsne_ctx *ctx);

/* Upload Y[N*no_dims] (row-major, as the reference) and reset momentum uY = 0 and gains = 1. */
int fitsne_set_Y(fitsne_ctx *ctx, const double *Y);
int fitsne_get_Y(fitsne_ctx *ctx, double *Y);

/* Gradient at the resident Y (no update).  dC[N*no_dims] (nullable) receives it; sum_Q (nullable)
 * receives current_sum_Q. */
int fitsne_gradient(fitsne_ctx *ctx, double *dC, double *sum_Q);

/* One iteration of TSNE::run's hot loop (tsne.cpp:440-531): gradient, gains/momentum/clip update,
 * Y += uY and zeroMean(Y).  Asynchronous on the context's stream; sum_Q (nullable) forces a sync.
 * The FP32 form's host-planned steps apply zeroMean as the next step reads Y, or before any other call
 * touches it (fitsne_get_Y, fitsne_kl, ...): every result is bitwise what an immediate zeroMean gives. */
int fitsne_step(fitsne_ctx *ctx, const fitsne_step_params *prm, double *sum_Q);

/* Captured steps (FP32 form; no reference counterpart).  When enabled, fitsne_step replays one hipGraph
 * holding the whole iteration with its grid geometry planned on the device (no host round trip per
 * step): results are bitwise those of the host-planned step.  Grids and buffers are sized for this
 * build's largest grid (2D FFT length 2048, 1D 4096); a step whose grid would exceed it is refused
 * on the device, leaves the embedding unchanged and makes the next synchronising call (fitsne_sync,
 * fitsne_get_Y, fitsne_kl, fitsne_gradient, fitsne_last_geometry, or fitsne_step with sum_Q) return
 * FITSNE_E_UNSUPPORTED.  Re-captured after set_P / set_Y / reorder. */
int fitsne_set_graph(fitsne_ctx *ctx, int enable);
/* Host-planned FP32 steps run the attractive term and the grid stages side by side on two streams, and in
 * 1D the convolution's kernel line on a third (enable = 1, the default); 0 runs every kernel of the step
 * on the context's stream, one after another (so that per-kernel timings are those of each kernel alone).  Results are the same either way.  No counterpart
 * in the reference (a tuning knob of this build). */
int fitsne_set_overlap(fitsne_ctx *ctx, int enable);

/* Autotuner (no reference counterpart: the reference's spread is a serial loop).  Times, on the
 * context's current Y and geometry, the FP32 spread's points per workgroup (512, 1024, 2048, 4096) and
 * the attractive term's row blocks (about N / 256, 1024, 2048 or 4096 rows, or blocks of up to 4096 rows
 * in 2D, 8192 in 1D, balanced by a cost model, one per CU; the windowed copy of P is rebuilt for each),
 * and keeps the fastest of each (a larger spread size only when more than 3% faster); every result is
 * bitwise the same for any choice.
 * spread_points (may be null) receives the spread's choice.  FP32 contexts with P and Y set.
 * fitsne_batch_create tunes each context's spread; a fused batch uses 4096-row attractive blocks. */
int fitsne_autotune(fitsne_ctx *ctx, int *spread_points);
/* The tuned sizes: points per spread workgroup and rows per attractive workgroup, -1 for balanced blocks
 * (nullable outputs). */
int fitsne_autotune_stats(fitsne_ctx *ctx, int *spread_points, int *attractive_rows);
/* Rows per attractive workgroup of an FP32 context (0: automatic, about N / 256; -1: blocks of up to 4096
 * rows (8192 in 1D) cut so that their modelled times match, one per CU; else a multiple of 64 up to 4096
 * (8192 in 1D)).  The forces are bitwise the same for every value. */
int fitsne_set_attractive_rows(fitsne_ctx *ctx, int rows);
/* Set the spread's points per workgroup directly (a multiple of 512 in [512, 8192]; default 1024). */
int fitsne_set_spread_points(fitsne_ctx *ctx, int points);

/* Batched embeddings (no reference counterpart): one optimisation step of every context of the batch
 * (FP32 contexts on one device, each with its own P and Y): each context's captured step is launched
 * on its own stream, so the embeddings' steps overlap on the device.  prm[n] holds each context's
 * parameters.  Results are bitwise those of stepping each context alone.  Creating a batch runs the
 * autotuner on each context. */
typedef struct fitsne_batch fitsne_batch;
int fitsne_batch_create(fitsne_batch **batch, fitsne_ctx **ctxs, int n);
int fitsne_batch_step(fitsne_batch *batch, const fitsne_step_params *prm);
int fitsne_batch_destroy(fitsne_batch *batch);

/* KL divergence at the resident Y with the Z of the last gradient (tsne.cpp:547-571 semantics:
 * stale Z, FLT_MIN guards, exaggerated P); the exaggeration correction is the caller's.  An exact
 * context evaluates evaluateError (tsne.cpp:1286-1326) with the Z of the current Y. */
int fitsne_kl(fitsne_ctx *ctx, double *C);

/* Wait for the context's stream. */
int fitsne_sync(fitsne_ctx *ctx);

/* Per-kernel timing: when enabled, each kernel launch on the context's stream is bracketed by
 * HIP events; fitsne_kernel_stats returns, for up to `cap` kernels, their names (NUL-separated in
 * names[cap*48]), launch counts and total milliseconds since the last reset. */
int fitsne_set_timing(fitsne_ctx *ctx, int enable);
int fitsne_kernel_stats(fitsne_ctx *ctx, int cap, char *names, int *counts, double *total_ms);
/* Geometry of the last gradient: out[0..5] = n_boxes_per_dim, n1d, fft_len, min_coord, max_coord,
 * nnz of P. */
int fitsne_last_geometry(fitsne_ctx *ctx, double *out);

/* Diagnostics (no reference counterpart): copy up to `bytes` of an internal device buffer of the last
 * gradient into dst; returns the number of bytes copied.  Names: "grid_in", "grid_out", "kq", "k1",
 * "ks", "kt", "zpart", "zbuf", "pos", "dC".  Layouts are those documented in the kernels. */
long long fitsne_debug_read(fitsne_ctx *ctx, const char *name, void *dst, long long bytes);

#ifdef __cplusplus
}
#endif
#endif
