/*
 * ref_run_harness.cpp -- one C entry point into the reference's OWN optimisation loop, TSNE::run
 * (FIt-SNE 1.2.1, tsne.cpp:118-592), compiled from the reference's sources by oracle/Makefile into
 * oracle/_ref/libfitsne_ref_run.so.  TEST INFRASTRUCTURE: it pins the oracle's restatement of the loop
 * (orc_run_bh / orc_run_exact, which share the loop of orc_run_fft).
 *
 * TSNE::run also contains the FFT gradient calls, which reach FFTW (absent from this image).  This
 * library is therefore linked with FFTW's five entry points left UNRESOLVED -- no definition of them,
 * and no other code standing in for FFTW, is linked -- and must be loaded with lazy binding
 * (RTLD_LAZY).  Only the Barnes-Hut (nbody_algorithm = 1) and exact (theta = 0) branches are driven, and
 * neither calls into FFTW; a call that did would stop the process with an unresolved-symbol error
 * instead of running substitute code.
 *
 * Affinities come from the files the reference's load_affinities = 1 branch reads (tsne.cpp:172-189:
 * P.dat; :233-276: P_row.dat, P_col.dat, P_val.dat) in the current directory, so the caller runs this in
 * a directory of its own.  The initial Y is given (skip_random_init).
 */
#include <cstdlib>

#include "nbodyfft.h"
#include "tsne.h"

#define API extern "C" __attribute__((visibility("default")))

API int ref_tsne_run(double *X, int N, int D, double *Y, int no_dims, double perplexity, double theta,
                     int max_iter, int stop_lying_iter, int mom_switch_iter, double momentum, double final_momentum,
                     double learning_rate, int nbody_algorithm, double early_exag_coeff, double *costs,
                     int no_momentum_during_exag, int start_late_exag_iter, double late_exag_coeff, double df,
                     double max_step_norm) {
    TSNE t;
    return t.run(X, N, D, Y, no_dims, perplexity, theta, /*rand_seed*/ -1, /*skip_random_init*/ true, max_iter,
                 stop_lying_iter, mom_switch_iter, momentum, final_momentum, learning_rate, /*K*/ -1, /*sigma*/ -1.0,
                 nbody_algorithm, /*knn_algo*/ 2, early_exag_coeff, costs, no_momentum_during_exag != 0,
                 start_late_exag_iter, late_exag_coeff, /*n_trees*/ 50, /*search_k*/ -1, /*nterms*/ 3,
                 /*intervals_per_integer*/ 1.0, /*min_num_intervals*/ 50, /*nthreads*/ 1, /*load_affinities*/ 1,
                 /*perplexity_list_length*/ 0, /*perplexity_list*/ nullptr, df, max_step_norm);
}
