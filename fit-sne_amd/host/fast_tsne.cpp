// fast_tsne -- the FIt-SNE command-line program on one MI355X.
//
// Same invocation, input and output as the reference program (tsne.cpp:2038-2136):
//     fast_tsne 1.2.1 [data_path=data.dat] [result_path=result.dat] [nthreads=0]
// reads the parameter/data file written by fast_tsne.py (layout of TSNE::load_data, tsne.cpp:1895-2022),
// runs the optimisation with the semantics of TSNE::run (tsne.cpp:118-592) and writes the result file
// (TSNE::save_data, tsne.cpp:2025-2036).  The gradient loop, the nearest-neighbour search, the
// perplexity calibration and the symmetrisation of P run on the GPU through libfitsne's C ABI; the host
// keeps the sums whose order defines the result: data centring and scaling, and the normalisation of P.
//
// Supported: nbody_algo = 2 (FFT interpolation) with theta != 0, no_dims 1 or 2; nbody_algo = 1
// (Barnes-Hut) with theta != 0, no_dims 1..3; and the exact mode (theta = 0, any nbody_algo, no_dims
// 1..4) with its dense P built on the GPU.  knn_algo 1 (Annoy) and 2 (VP tree) both use the exact
// neighbour search (Annoy's approximate search is not reproduced).  FITSNE_PRECISION=fp64 in the
// environment selects the FP64 plain GPU path of the FFT mode (default: the FP32 form); the exact and
// Barnes-Hut modes are FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "fitsne_abi.h"

namespace {

const char kVersion[] = "1.2.1";

struct Params {
    int n = 0, d = 0, no_dims = 2, max_iter = 1000, stop_lying_iter = 250, mom_switch_iter = 250;
    double theta = 0.5, perplexity = 30.0;
    std::vector<double> perplexity_list;
    double momentum = 0.5, final_momentum = 0.8, learning_rate = 200.0, max_step_norm = 5.0;
    int K = -1;
    double sigma = -1.0;
    int nbody_algo = 2, knn_algo = 2;
    double early_exag_coeff = 12.0;
    int no_momentum_during_exag = 0, n_trees = 50, search_k = -1, start_late_exag_iter = -1;
    double late_exag_coeff = -1.0;
    int nterms = 3;
    double intervals_per_integer = 1.0;
    int min_num_intervals = 50;
    int rand_seed = 0;
    double df = 1.0;
    int load_affinities = 0;
};

[[noreturn]] void die(const std::string &msg, int code = 1) {
    std::fprintf(stderr, "fast_tsne: %s\n", msg.c_str());
    std::exit(code);
}

void check(int rc, const char *what) {
    if (rc != FITSNE_OK) die(std::string(what) + ": " + fitsne_last_error());
}

template <typename T>
bool get(FILE *f, T *v, size_t n = 1) {
    return std::fread(v, sizeof(T), n, f) == n;
}

// Field order of tsne.cpp:1915-1985.  The optional tail (seed, df, load_affinities, initial Y) is read
// while the file has bytes left; Y counts only if it is complete (tsne.cpp:1976-1985).
bool load_data(const char *path, Params &p, std::vector<double> &X, std::vector<double> &Y, bool &have_init) {
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        std::printf("[fast_tsne] cannot open the data file %s\n", path);
        return false;
    }
    bool ok = get(f, &p.n) && get(f, &p.d) && get(f, &p.theta) && get(f, &p.perplexity);
    if (ok && p.perplexity == 0) {
        int L = 0;
        ok = get(f, &L) && L > 0;
        if (ok) {
            p.perplexity_list.resize(L);
            ok = get(f, p.perplexity_list.data(), (size_t)L);
        }
    }
    ok = ok && get(f, &p.no_dims) && get(f, &p.max_iter) && get(f, &p.stop_lying_iter) && get(f, &p.mom_switch_iter) &&
         get(f, &p.momentum) && get(f, &p.final_momentum) && get(f, &p.learning_rate) && get(f, &p.max_step_norm) &&
         get(f, &p.K) && get(f, &p.sigma) && get(f, &p.nbody_algo) && get(f, &p.knn_algo) &&
         get(f, &p.early_exag_coeff) && get(f, &p.no_momentum_during_exag) && get(f, &p.n_trees) &&
         get(f, &p.search_k) && get(f, &p.start_late_exag_iter) && get(f, &p.late_exag_coeff) && get(f, &p.nterms) &&
         get(f, &p.intervals_per_integer) && get(f, &p.min_num_intervals);
    if (!ok || p.n <= 0 || p.d <= 0 || p.no_dims <= 0 || p.max_iter < 0) {
        std::fclose(f);
        die("malformed data file header");
    }
    if (p.nbody_algo == 2 && p.no_dims > 2)
        die("the FFT gradient needs no_dims 1 or 2, got " + std::to_string(p.no_dims));
    X.resize((size_t)p.n * p.d);
    if (!get(f, X.data(), X.size())) {
        std::fclose(f);
        die("data file ends inside the data matrix");
    }
    if (get(f, &p.rand_seed))
        if (get(f, &p.df)) (void)get(f, &p.load_affinities);
    Y.assign((size_t)p.n * p.no_dims, 0.0);
    have_init = get(f, Y.data(), Y.size());
    std::fclose(f);
    // one "name=value" pair per setting, in file order (nothing parses this; it is a log)
    std::printf("[fast_tsne] data: N=%d D=%d\n", p.n, p.d);
    std::printf("[fast_tsne] settings: theta=%g perplexity=%g no_dims=%d max_iter=%d stop_lying_iter=%d "
                "mom_switch_iter=%d momentum=%g final_momentum=%g learning_rate=%g max_step_norm=%g K=%d sigma=%g "
                "nbody_algo=%d knn_algo=%d early_exag_coeff=%g no_momentum_during_exag=%d n_trees=%d search_k=%d "
                "start_late_exag_iter=%d late_exag_coeff=%g nterms=%d intervals_per_integer=%g "
                "min_num_intervals=%d df=%g seed=%d load_affinities=%d\n",
                p.theta, p.perplexity, p.no_dims, p.max_iter, p.stop_lying_iter, p.mom_switch_iter, p.momentum,
                p.final_momentum, p.learning_rate, p.max_step_norm, p.K, p.sigma, p.nbody_algo, p.knn_algo,
                p.early_exag_coeff, p.no_momentum_during_exag, p.n_trees, p.search_k, p.start_late_exag_iter,
                p.late_exag_coeff, p.nterms, p.intervals_per_integer, p.min_num_intervals, p.df, p.rand_seed,
                p.load_affinities);
    if (!p.perplexity_list.empty()) {
        std::printf("[fast_tsne] perplexities:");
        for (double v : p.perplexity_list) std::printf(" %g", v);
        std::printf("\n");
    }
    std::printf("[fast_tsne] initial embedding: %s\n", have_init ? "given" : "random");
    return true;
}

void save_data(const char *path, const std::vector<double> &Y, const std::vector<double> &costs, int n, int d,
               int max_iter) {
    FILE *f = std::fopen(path, "wb");
    if (!f) die("could not open the result file");
    std::fwrite(&n, sizeof(int), 1, f);
    std::fwrite(&d, sizeof(int), 1, f);
    std::fwrite(Y.data(), sizeof(double), Y.size(), f);
    std::fwrite(&max_iter, sizeof(int), 1, f);
    std::fwrite(costs.data(), sizeof(double), costs.size(), f);
    std::fclose(f);
    std::printf("[fast_tsne] result: %d x %d embedding written to %s\n", n, d, path);
}

// Column means subtracted, each mean summed over the rows in order then divided by N (tsne.cpp:1851-1876).
void zero_mean(std::vector<double> &X, int N, int D) {
    std::vector<double> mean(D, 0.0);
    for (int n = 0; n < N; n++)
        for (int d = 0; d < D; d++) mean[d] += X[(size_t)n * D + d];
    for (int d = 0; d < D; d++) mean[d] /= (double)N;
    for (int n = 0; n < N; n++)
        for (int d = 0; d < D; d++) X[(size_t)n * D + d] -= mean[d];
}

// Marsaglia's polar method on the C library's rand() (the generator TSNE::randn uses, tsne.cpp:1878-1890).
double randn() {
    double x, y, r;
    do {
        x = 2 * (std::rand() / ((double)RAND_MAX + 1)) - 1;
        y = 2 * (std::rand() / ((double)RAND_MAX + 1)) - 1;
        r = x * x + y * y;
    } while (r >= 1.0 || r == 0.0);
    return x * std::sqrt(-2 * std::log(r) / r);
}

struct Csr {
    std::vector<uint32_t> row, col;
    std::vector<double> val;
};

// (P + P^T) / 2 with the entry order and values of the reference's symmetrizeMatrix
// (tsne.cpp:1730-1828), computed on the GPU.
Csr symmetrize(const Csr &P, int N) {
    Csr S;
    const size_t cap = std::max<size_t>(2 * P.col.size(), 1);
    S.row.resize((size_t)N + 1);
    S.col.resize(cap);
    S.val.resize(cap);
    long long m = 0;
    check(fitsne_symmetrize(P.row.data(), P.col.data(), P.val.data(), N, S.row.data(), S.col.data(), S.val.data(), &m),
          "symmetrisation");
    S.col.resize((size_t)m);
    S.val.resize((size_t)m);
    return S;
}

template <typename T>
bool read_file(const char *name, std::vector<T> &v, size_t n) {
    FILE *f = std::fopen(name, "rb");
    if (!f) return false;
    v.resize(n);
    const bool ok = std::fread(v.data(), sizeof(T), n, f) == n;
    std::fclose(f);
    return ok;
}

template <typename T>
void write_file(const char *name, const std::vector<T> &v) {
    FILE *f = std::fopen(name, "wb");
    if (!f) die(std::string("could not write ") + name);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// The data as TSNE::run prepares it before the similarities (tsne.cpp:153-165): centred, and scaled
// by the largest absolute value when a perplexity is used.
void prepare_data(const Params &p, std::vector<double> &X) {
    zero_mean(X, p.n, p.d);
    if (p.perplexity > 0 || !p.perplexity_list.empty()) {
        std::printf("[fast_tsne] data centred and divided by max |x| (perplexity calibration)\n");
        double mx = 0.0;
        for (double v : X) mx = std::max(mx, std::fabs(v));
        for (double &v : X) v /= mx;
    } else {
        std::printf("[fast_tsne] data centred, not rescaled (fixed sigma)\n");
    }
}

// Exact mode (theta = 0, tsne.cpp:171-229): the dense P, built on the device by
// fitsne_exact_affinities (calibrated rows, P + P^T, divided by the total) or read from P.dat, stays in
// the context.  Returns the largest row sum of P when asked (the automatic early exaggeration).
double exact_affinities(Params &p, std::vector<double> &X, fitsne_ctx *ctx, bool want_max_row) {
    const int N = p.n, D = p.d;
    const size_t nn = (size_t)N * N;
    std::vector<double> P;
    if (p.load_affinities == 1) {
        std::printf("[fast_tsne] affinities: dense P read from P.dat\n");
        if (!read_file("P.dat", P, nn)) die("could not read P.dat");
        check(fitsne_set_P_dense(ctx, N, P.data()), "upload P");
    } else {
        prepare_data(p, X);
        std::printf("[fast_tsne] affinities: dense P on the GPU (theta = 0)\n");
        const bool need_host = p.load_affinities == 2 || want_max_row;
        if (need_host) P.resize(nn);
        double sum_P = 0.0;
        const double *plist = p.perplexity_list.empty() ? nullptr : p.perplexity_list.data();
        check(fitsne_exact_affinities(ctx, X.data(), N, D, p.perplexity, p.sigma, (int)p.perplexity_list.size(), plist,
                                      need_host ? P.data() : nullptr, &sum_P),
              "exact input similarities");
        std::printf("[fast_tsne] dense P total before normalisation: %.17g\n", sum_P);
        if (p.load_affinities == 2) {
            std::printf("[fast_tsne] dense P written to P.dat\n");
            write_file("P.dat", P);
        }
    }
    double mx = 0.0;
    if (want_max_row)
        for (int n = 0; n < N; n++) {
            double s = 0.0;
            for (int m = 0; m < N; m++) s += P[(size_t)n * N + m];
            mx = std::max(mx, s);
        }
    return mx;
}

// Input similarities: kNN rows calibrated by perplexity (or a fixed sigma), symmetrised, normalised
// (tsne.cpp:151-366).
Csr affinities(Params &p, std::vector<double> &X) {
    const int N = p.n, D = p.d;
    Csr P;
    if (p.load_affinities == 1) {
        std::printf("[fast_tsne] affinities: sparse P read from P_row.dat, P_col.dat, P_val.dat\n");
        if (!read_file("P_row.dat", P.row, (size_t)N + 1)) die("could not read P_row.dat");
        const size_t nnz = P.row[N];
        if (!read_file("P_val.dat", P.val, nnz) || !read_file("P_col.dat", P.col, nnz))
            die("could not read P_val.dat / P_col.dat");
        return P;
    }
    prepare_data(p, X);
    int K;
    double sigma;
    if (p.perplexity < 0) {
        std::printf("[fast_tsne] affinities: fixed sigma %g with K = %d\n", p.sigma, p.K);
        K = p.K;
        sigma = p.sigma;
    } else {
        if (p.perplexity > 0) {
            K = (int)(3 * p.perplexity);
        } else {
            K = (int)(3 * p.perplexity_list[0]);
            for (size_t m = 1; m < p.perplexity_list.size(); m++)
                if (3 * p.perplexity_list[m] > K) K = (int)(3 * p.perplexity_list[m]);
        }
        sigma = -1.0;
    }
    if (K < 1 || K > N - 1) die("the number of neighbours K = " + std::to_string(K) + " must be in [1, N-1]");
    if (p.perplexity > K) std::printf("[fast_tsne] warning: perplexity %g exceeds K = %d\n", p.perplexity, K);
    std::printf("[fast_tsne] exact %d nearest neighbours on the GPU\n", K);
    std::vector<int32_t> idx((size_t)N * K);
    std::vector<double> dist((size_t)N * K);
    check(fitsne_knn_exact(X.data(), N, D, K, idx.data(), dist.data()), "nearest neighbours");
    // distances2similarities list overload: the distances are used as they are (tsne.cpp:1474-1500)
    std::vector<double> row((size_t)N * K), beta(N);
    const std::vector<double> single = {p.perplexity};
    const std::vector<double> &plist = p.perplexity != 0 ? single : p.perplexity_list;
    std::vector<double> tmp((size_t)N * K);
    for (size_t m = 0; m < plist.size(); m++) {
        check(fitsne_distances2similarities_batch(dist.data(), m == 0 ? row.data() : tmp.data(), N, K, plist[m], sigma,
                                                  1, nullptr, beta.data(), nullptr),
              "perplexity calibration");
        if (m > 0)
            for (size_t i = 0; i < row.size(); i++) row[i] += tmp[i];
    }
    if (plist.size() > 1)
        for (double &v : row) v /= (double)plist.size();
    P.row.resize((size_t)N + 1);
    for (int n = 0; n <= N; n++) P.row[n] = (uint32_t)((size_t)n * K);
    P.col.assign(idx.begin(), idx.end());
    P.val = row;
    std::printf("[fast_tsne] symmetrising P on the GPU\n");
    Csr S = symmetrize(P, N);
    double sum = 0.0;
    for (double v : S.val) sum += v;
    for (double &v : S.val) v /= sum;
    if (p.load_affinities == 2) {
        std::printf("[fast_tsne] sparse P written to P_row.dat, P_col.dat, P_val.dat\n");
        write_file("P_val.dat", S.val);
        write_file("P_col.dat", S.col);
        write_file("P_row.dat", S.row);
    }
    return S;
}

}  // namespace

int main(int argc, char **argv) {
    std::printf("[fast_tsne] FIt-SNE %s interface on MI355X\n", kVersion);
    if (argc < 2) {
        std::printf("[fast_tsne] usage: fast_tsne %s [data_path] [result_path] [nthreads]\n", kVersion);
        return -1;
    }
    if (std::strcmp(argv[1], kVersion) != 0) {
        std::printf("[fast_tsne] wrong version %s: this program speaks %s\n", argv[1], kVersion);
        return -1;
    }
    const char *data_path = argc >= 3 ? argv[2] : "data.dat";
    const char *result_path = argc >= 4 ? argv[3] : "result.dat";
    const unsigned nthreads = argc >= 5 ? (unsigned)std::strtoul(argv[4], nullptr, 10) : 0u;
    std::printf("[fast_tsne] data %s, result %s, nthreads %u (unused: the GPU sets the parallelism)\n", data_path,
                result_path, nthreads);

    Params p;
    std::vector<double> X, Y;
    bool have_init = false;
    if (!load_data(data_path, p, X, Y, have_init)) return 0;
    const int N = p.n, nd = p.no_dims;
    if (N - 1 < 3 * p.perplexity) die("perplexity too large: 3 x perplexity must be below N");
    const bool exact = p.theta == 0.0;
    const bool bh = !exact && p.nbody_algo == 1;
    if (!exact && p.nbody_algo != 1 && p.nbody_algo != 2) die("undefined algorithm: nbody_algo must be 1 or 2");
    if (!exact && p.knn_algo != 1 && p.knn_algo != 2) die("invalid knn_algo: must be 1 or 2");
    if (exact && nd > 4) die("the exact mode of this build supports no_dims <= 4");
    if (bh && nd > 3) die("the Barnes-Hut mode of this build supports no_dims <= 3");
    if (fitsne_device_count() < 1) die("no GPU device is visible");

    fitsne_ctx *ctx = nullptr;
    double max_row_sum = 0.0;
    const char *prec = std::getenv("FITSNE_PRECISION");
    const int precision = exact ? FITSNE_FP64_EXACT
                          : bh ? FITSNE_FP64_BH
                          : (prec && std::strcmp(prec, "fp64") == 0) ? FITSNE_FP64_PLAIN : FITSNE_FP32;
    if (exact) {
        check(fitsne_ctx_create(&ctx, 0, precision, nd, p.nterms, p.intervals_per_integer, p.min_num_intervals, p.df),
              "context");
        max_row_sum = exact_affinities(p, X, ctx, p.early_exag_coeff == 0);
        std::printf("[fast_tsne] affinities ready (dense)\n");
    } else {
        Csr P = affinities(p, X);
        std::printf("[fast_tsne] affinities ready: %u entries, density %.3g\n", P.row[N],
                    (double)P.row[N] / ((double)N * (double)N));
        if (p.early_exag_coeff == 0)
            for (int n = 0; n < N; n++) {
                double s = 0.0;
                for (uint32_t k = P.row[n]; k < P.row[n + 1]; k++) s += P.val[k];
                max_row_sum = std::max(max_row_sum, s);
            }
        check(fitsne_ctx_create(&ctx, 0, precision, nd, p.nterms, p.intervals_per_integer, p.min_num_intervals, p.df),
              "context");
        check(fitsne_set_P(ctx, N, P.row.data(), P.col.data(), P.val.data()), "upload P");
        if (bh) {
            std::printf("[fast_tsne] Barnes-Hut gradient, theta %g\n", p.theta);
            check(fitsne_set_theta(ctx, p.theta), "theta");
        }
    }

    if (!have_init) {
        if (p.rand_seed >= 0) {
            std::printf("[fast_tsne] random start, seed %d\n", p.rand_seed);
            std::srand((unsigned)p.rand_seed);
        } else {
            std::printf("[fast_tsne] random start, seeded from the clock\n");
            std::srand((unsigned)std::time(nullptr));
        }
        for (double &v : Y) v = randn() * .0001;
    }

    // largest stable exaggeration (tsne.cpp:389-401; the reference's exact mode has no row_P there, the
    // row sums of the dense P are its natural counterpart)
    double early = p.early_exag_coeff;
    if (early == 0) early = 1.0 / (p.learning_rate * max_row_sum);
    std::printf("[fast_tsne] early exaggeration %g\n", early);

    double exag = early;
    check(fitsne_set_exaggeration(ctx, exag), "exaggeration");
    check(fitsne_set_Y(ctx, Y.data()), "upload Y");
    std::printf("[fast_tsne] optimising (%s)\n", precision == FITSNE_FP32        ? "FP32 form"
                                                 : precision == FITSNE_FP64_EXACT ? "exact, FP64"
                                                 : precision == FITSNE_FP64_BH    ? "Barnes-Hut, FP64"
                                                                                  : "FP64 plain");

    std::vector<double> costs((size_t)p.max_iter, 0.0);
    double momentum = p.momentum;
    auto t0 = std::chrono::steady_clock::now();
    for (int iter = 0; iter < p.max_iter; iter++) {
        fitsne_step_params sp{};
        sp.momentum = momentum;
        sp.learning_rate = p.learning_rate;
        // no_momentum_during_exag: plain descent up to stop_lying_iter, then gains/momentum without the
        // step clip (tsne.cpp:478-490); otherwise gains/momentum with the clip (tsne.cpp:491-514)
        sp.plain_descent = p.no_momentum_during_exag && iter <= p.stop_lying_iter;
        sp.max_step_norm = p.no_momentum_during_exag ? 0.0 : p.max_step_norm;
        check(fitsne_step(ctx, &sp, nullptr), "gradient step");
        bool changed = false;
        if (iter == p.stop_lying_iter) {
            std::printf("[fast_tsne] iteration %d: early exaggeration off\n", iter);
            exag /= early;
            changed = true;
        }
        if (iter == p.start_late_exag_iter) {
            std::printf("[fast_tsne] iteration %d: late exaggeration %g\n", iter, p.late_exag_coeff);
            exag *= p.late_exag_coeff;
            changed = true;
        }
        if (changed) check(fitsne_set_exaggeration(ctx, exag), "exaggeration");
        if (iter == p.mom_switch_iter) momentum = p.final_momentum;
        if ((iter + 1) % 50 == 0 || iter == p.max_iter - 1) {
            double C = 0.0;
            check(fitsne_kl(ctx, &C), "KL divergence");
            if (iter < p.stop_lying_iter && p.stop_lying_iter != -1) C = C / early - std::log(early);
            if (iter >= p.start_late_exag_iter && p.start_late_exag_iter != -1)
                C = C / p.late_exag_coeff - std::log(p.late_exag_coeff);
            costs[iter] = C;
            const auto now = std::chrono::steady_clock::now();
            std::printf("[fast_tsne] iteration %d: KL %.6f (%.2f s since the last report)\n", iter + 1, C,
                        std::chrono::duration<double>(now - t0).count());
            t0 = now;
        }
    }
    check(fitsne_get_Y(ctx, Y.data()), "download Y");
    fitsne_ctx_destroy(ctx);
    save_data(result_path, Y, costs, N, nd, p.max_iter);
    std::printf("[fast_tsne] done\n");
    return 0;
}
