// fitsne_host.hip -- host side of libfitsne.so: contexts, FFT plans, per-iteration geometry, the FP64
// plain and FP32 gradient pipelines, the mirror tier and the C-ABI of include/fitsne_abi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fitsne_abi.h"
#include "fitsne_kernels.hpp"
#include "fitsne_batch.hpp"
#include "fitsne_exact.hpp"
#include "fitsne_sym.hpp"
#include "fitsne_bh.hpp"
#include "fitsne_knn.hpp"
#include "fitsne_knn_mfma.hpp"
#include "fitsne_mlp.hpp"

using namespace fitsne;

// ============================================================================ errors

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPTRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FITSNE_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int r_ = (expr);           \
        if (r_ != FITSNE_OK) return r_; \
    } while (0)

// ============================================================================ limits

constexpr int kMaxTot32 = 8192;   // complex FP32 elements per FFT block (64 KB LDS)
constexpr int kMaxTot64 = 4096;   // complex FP64 elements per FFT block (64 KB LDS)
constexpr int kConv1dThreads = 1024;
// 2D FFT kernels: threads per block and the bound on (lines x length) each is compiled for; the FFT
// length of a 2D grid is limited to kMaxLen2d (n1d <= 1024 nodes per dimension).
constexpr int kMaxLen2d = 2048;
constexpr int kKernThreads = 256, kKernTot = kMaxLen2d;        // kernel spectra: 1 line per block
constexpr int kRowThreads = 256, kRowTot = 2 * kMaxLen2d;      // row passes: 2 packed lines
constexpr int kColThreads = 512, kColTot32 = 4 * kMaxLen2d;   // FP32 column pass: 4 lines
constexpr int kColThreads64 = 512, kColTot64 = 2 * kMaxLen2d;  // FP64 plain column pass

// ============================================================================ device buffers

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    template <typename T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

static int ensure(DBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (bytes <= b.cap) return FITSNE_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 8;  // headroom: grids change size every iteration
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        return fail(FITSNE_E_ALLOC, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    b.cap = want;
    return FITSNE_OK;
}

static void release(DBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// FFT plans and the geometry rules (fft_size, make_plan, n_boxes_2d / _1d, make_geo) live in
// fitsne_common.hpp, shared with the device.

// ============================================================================ context

struct KStat {
    int count = 0;
    double ms = 0;
};

struct fitsne_ctx {
    int device = 0;
    int precision = FITSNE_FP32;
    int D = 2, p = 3, mni = 50;
    double ipi = 1.0, df = 1.0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;  // the context's stream; `stream` is a batch's while it is a member
    struct fitsne_batch *batch = nullptr;  // the fused batch this context belongs to
    long layout_gen = 0;  // bumped whenever device buffers or launch shapes may have changed
    // second stream of the host-planned FP32 step: the attractive term runs on it beside the grid stages
    hipStream_t astream = nullptr;
    int spread_pts = kSpreadPts;  // points per spread workgroup (fitsne_autotune); results do not depend on it
    bool overlap = true;          // host-planned FP32 step: attractive term on astream beside the grid stages
    int att_rows = 0;             // rows per attractive workgroup (0: about N / 256); results do not depend on it
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // kernel spectra on astream: after the geometry record or a new twiddle table (when the main stream
    // wrote either in this step), before the column pass
    hipEvent_t geo_ev = nullptr, kern_ev = nullptr;
    // 1D host-planned FP32 step: the kernel line's spectra on a third stream, beside the attractive kernel
    // and the spread, into kline[3][Mp]; kl_go orders them after the main stream's work so far, kl_ev
    // orders the convolution after them
    hipStream_t kstream = nullptr;
    hipEvent_t kl_go = nullptr, kl_ev = nullptr;
    // 1D host-planned FP32 step: the attractive kernel stays on the main stream and the rest of the step
    // is queued on astream until the interpolation, which returns to main (swap_main: main meanwhile)
    hipStream_t swap_main = nullptr;
    bool join_pending = false;

    int N = 0;
    long long nnz = 0;
    double exag = 1.0;  // current multiplier on P
    // P
    DBuf rowP, colP, val;  // rowP int64[N+1], colP uint32[nnz], val T[nnz]
    // embedding state (element type = float for FP32, double for FP64 plain)
    DBuf Y, uY, gains, pos, dC;
    DBuf kline;
    // sort (FP32)
    // exact kNN (affinity build)
    DBuf knnX, knnXT, knnI, knnD;
    // windowed copy of P for the FP32 attractive term (k_attr_win), built on first use
    DBuf wcount, wstart, wpure, wkey, wval, wbmax, wrstart;
    bool win_ok = false;
    int win_R = 0, win_NB = 0, win_NW = 0;
    bool win_bal = false;  // row blocks from balance_blocks (starts in wrstart)
    // grids, spectra
    DBuf gin, gout, spec, ospec, Trow, KQ, K1, KS, KT, zpart, zbuf, part, mean, mmpart, mm, fixgrid;
    DBuf mmu;      // per block and coordinate (max, min) of the updated Y, from interp_f32 for mean_mm
    // device records of the step (geometry and plan; optimiser parameters) and the captured step
    DBuf dgeo, dstep, dtw;
    DBuf dlive;                  // a geometry record with live = 1, for launches before the step's geometry
    DBuf klpart;                 // KL block partials (never a buffer a captured step uses)
    bool dlive_set = false;
    long graph_gen = 0;          // incremented at every capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool use_graph = false;      // fitsne_step replays a captured step (FP32 form)
    bool mm_parts_valid = false; // mmpart holds the min / max partials of the resident Y
    // FP32 host-planned steps leave zeroMean pending: Y in memory lacks the last mean (in `mean`), and
    // the next step's readers (attractive, spread, interpolation, min rule) apply it on load; anything
    // else that touches Y applies it first (materialize)
    bool ctr_lazy = false;
    bool tw_fresh = false;  // a twiddle table was written on the stream since the last step's stages
    bool graph_ran = false;      // a captured step ran since the last error check
    // plain path extras
    DBuf xs, ys, box, wx, wy, ch, w, V, potA, potB;
    // exact path (dense P in val): per-row attractive / repulsive sums and Z shares
    DBuf att, rep, zrow;
    // Barnes-Hut: the host tree of the current Y and its device copy (pre-order), theta
    double theta = 0.5;
    BhTree<1> bh_tree1;
    BhTree<2> bh_tree2;
    BhTree<3> bh_tree3;
    std::vector<double> bh_y;
    void *bh_pin = nullptr;  // pinned host memory: Y, then the flattened tree
    size_t bh_pin_cap = 0;
    size_t bh_nodes = 0;
    DBuf bh_tree_d, bh_neg, bh_Q;
    // symmetrisation of a CSR P (mirror tier)
    DBuf sy_row, sy_col, sy_val, sy_erow, sy_keys, sy_keys2, sy_vals, sy_vals2, sy_partner, sy_cnt, sy_start,
        sy_okeys, sy_okeys2, sy_tmp, sy_scol, sy_sval;
    // twiddles per FFT length
    std::map<int, DBuf> tw32, tw64;
    // min/max handoff from the last step
    double *h_mm = nullptr;      // pinned [4]
    double *h_mm_dev = nullptr;  // h_mm as the device addresses it
    hipEvent_t mm_event = nullptr;
    bool mm_pending = false, mm_valid = false;
    double mm_min = 0, mm_max = 0;
    double mm_first = 0;  // the caller's first coordinate (xs[0]) at the last min / max
    // kernel-spectrum cache key
    bool kcache = false;
    double kc_mn = 0, kc_mx = 0;
    int kc_nb = 0;
    // last geometry
    Geo last{};
    double last_min = 0, last_max = 0;
    double last_sumQ = 0;
    bool have_grad = false;
    // spatial order of the points (FP32 form): internal row r holds the caller's point perm[r]
    bool reorder_on_set_Y = true;
    std::vector<int> perm, inv;         // internal -> caller, caller -> internal
    std::vector<int64_t> rowP_host;     // internal-order row offsets
    DBuf inv_d, colP2, val2, tmp;       // device copy of inv; double buffers for permutation
    bool permuted = false;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pend {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Pend> pend;
    std::map<std::string, KStat> stats;
};

static hipEvent_t take_event(fitsne_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <typename... KA, typename... A>
static void launch(fitsne_ctx *c, const char *name, void (*k)(KA...), dim3 grid, dim3 block, size_t shm,
                   A... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c->timing) {
        a = take_event(c);
        b = take_event(c);
        if (a) (void)hipEventRecord(a, c->stream);
    }
    k<<<grid, block, shm, c->stream>>>(args...);
    if (c->timing && a && b) {
        (void)hipEventRecord(b, c->stream);
        c->pend.push_back({name, a, b});
    }
}

static int flush_timing(fitsne_ctx *c) {
    if (c->pend.empty()) return FITSNE_OK;
    HIPTRY(hipStreamSynchronize(c->stream));
    for (auto &p : c->pend) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        KStat &s = c->stats[p.name];
        s.count++;
        s.ms += ms;
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->pend.clear();
    return FITSNE_OK;
}

static int sync_ctx(fitsne_ctx *c) {
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(c->stream));
    if (c->astream) HIPTRY(hipStreamSynchronize(c->astream));
    if (c->kstream) HIPTRY(hipStreamSynchronize(c->kstream));
    return flush_timing(c);
}

static void drop_graph(fitsne_ctx *c);
static int check_graph(fitsne_ctx *c);
static int prepare_captured(fitsne_ctx *c);

template <typename T>
static int twiddles(fitsne_ctx *c, int n, const cx<T> **out) {
    auto &m = std::is_same<T, float>::value ? c->tw32 : c->tw64;
    auto it = m.find(n);
    if (it == m.end()) {
        DBuf b;
        TRY(ensure(b, sizeof(cx<T>) * (size_t)n));
        launch(c, "twiddles", k_twiddles<T>, dim3((n + 255) / 256), dim3(256), 0, b.as<cx<T>>(), n);
        c->tw_fresh = true;  // written on c->stream just now: a reader on another stream must wait
        it = m.emplace(n, b).first;
    }
    *out = it->second.template as<cx<T>>();
    return FITSNE_OK;
}

static int blocks_for(long n, int per = kThreads) { return (int)((n + per - 1) / per); }

// Workgroups of the passes over Y that leave per-block min / max partials (centring, min / max, and
// the captured and batched geometry that reduces them): up to 2048 coordinates each, at most 1024.
// Every writer and reader of a context's partials uses this count.  (With one coordinate per thread,
// the 64-problem batch's centring ran ~25,000 small workgroups.)
static int part_blocks(long n) { return std::min(blocks_for(n, 8 * kThreads), 1024); }

// Number of bits that hold every value in [0, v].
static int bits_for(unsigned long long v) {
    int b = 1;
    while (b < 64 && (1ull << b) <= v) b++;
    return b;
}

// lanes per CSR row: each lane carries 4 entries per pass, so 32 lanes cover 128 entries at once
static int sw_for(long long nnz, int N) {
    const double avg = N > 0 ? (double)nnz / N : 0;
    return avg >= 64 ? 32 : 16;
}

// ---------------------------------------------------------------- min / max of the resident Y

// Flat index (internal layout) of the caller's first coordinate, xs[0]: the value the 2D `else if`
// rule never tests against the running min when it opens the sequence.
static long first_coord(const fitsne_ctx *c) { return (long)(c->inv.empty() ? 0 : c->inv[0]) * c->D; }
static long skip_coord(const fitsne_ctx *c) { return c->D == 2 ? first_coord(c) : -1; }

template <typename TV>
static int minmax_now(fitsne_ctx *c, double *mn, double *mx) {
    const long n = (long)c->N * c->D;
    const int nblk = part_blocks(n);
    TRY(ensure(c->mmpart, sizeof(double) * 2 * nblk));
    TRY(ensure(c->mm, sizeof(double) * 4));
    launch(c, "minmax_partial", k_minmax_partial<TV>, dim3(nblk), dim3(kThreads), 0, c->Y.as<const TV>(), n,
           skip_coord(c), c->mmpart.as<double>());
    launch(c, "minmax_final", k_minmax_final<TV>, dim3(1), dim3(kWave), 0, c->mmpart.as<const double>(), nblk,
           c->Y.as<const TV>(), first_coord(c), c->mm.as<double>());
    HIPTRY(hipMemcpyAsync(c->h_mm, c->mm.p, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    *mx = c->h_mm[0];
    *mn = c->h_mm[1];
    return FITSNE_OK;
}

// Resolve the reference's 2D `else if` min (tsne.cpp:1045-1048): equal to min(rest) unless Y[0] is the
// strict minimum; then replay the sequential rule on the device.
template <typename TV>
static int resolve_quirk(fitsne_ctx *c, double y0, double *mn, double *mx) {
    if (c->D != 2 || !(y0 < *mn)) return FITSNE_OK;
    launch(c, "minmax_seq", k_minmax_seq<TV>, dim3(1), dim3(1024), 0, c->Y.as<const TV>(),
           c->permuted ? c->inv_d.as<const int>() : (const int *)nullptr, (long)c->N * 2, c->mm.as<double>(),
           c->ctr_lazy ? c->mean.as<const double>() : (const double *)nullptr);
    HIPTRY(hipMemcpyAsync(c->h_mm, c->mm.p, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    *mx = c->h_mm[0];
    *mn = c->h_mm[1];
    return FITSNE_OK;
}

template <typename TV>
static int current_minmax(fitsne_ctx *c, double *mn, double *mx) {
    if (c->mm_pending) {
        HIPTRY(hipEventSynchronize(c->mm_event));
        c->mm_pending = false;
        c->mm_valid = true;
        c->mm_max = c->h_mm[0];
        c->mm_min = c->h_mm[1];
        double y0 = c->h_mm[2];
        c->mm_first = y0;
        TRY(resolve_quirk<TV>(c, y0, &c->mm_min, &c->mm_max));
    }
    if (!c->mm_valid) {
        TRY(minmax_now<TV>(c, &c->mm_min, &c->mm_max));
        double y0 = c->h_mm[2];
        c->mm_first = y0;
        TRY(resolve_quirk<TV>(c, y0, &c->mm_min, &c->mm_max));
        c->mm_valid = true;
    }
    *mn = c->mm_min;
    *mx = c->mm_max;
    return FITSNE_OK;
}

static int geometry(fitsne_ctx *c, double mn, double mx, Geo *g) {
    // Y is finite on upload; the bounds are not when the optimiser has diverged, or in 2D when the
    // coordinates increase strictly in input order: the reference's `else if` rule (tsne.cpp:1045-1048)
    // then never assigns its minimum and its grid would be built from an infinite span.
    if (!(std::isfinite(mn) && std::isfinite(mx)))
        return fail(FITSNE_E_ARG, "grid bounds are not finite: non-finite coordinates, or (2D) coordinates that "
                                  "increase strictly in input order, which leave the min rule without a minimum");
    const int nb = c->D == 2 ? n_boxes_2d(mn, mx, c->ipi, c->mni) : n_boxes_1d(mn, mx, c->ipi, c->mni);
    if (nb < 1) return fail(FITSNE_E_ARG, "n_boxes < 1");
    *g = make_geo(c->D, c->p, mn, mx, nb);
    const bool f32 = c->precision == FITSNE_FP32;
    if (f32) {
        // the first x coordinate lies below the grid when the 2D min rule skipped it (extrapolated)
        const double gap = c->D == 2 && c->mm_first < mn ? (mn - c->mm_first) / g->bw : 0.0;
        const double lb = lagrange_inbox_bound(c->p);
        g->fix = spread_fix_scale(*g, (double)c->N, gap, lb);
        g->fix_fast = spread_fix_fast(*g, gap, lb);
        if (g->fix == 0)
            return fail(FITSNE_E_UNSUPPORTED,
                        "the first point lies " + std::to_string(gap) + " box widths below the grid (the 2D min "
                        "rule leaves it out); its extrapolated spread weights exceed the FP32 form's "
                        "fixed-point range: use the FP64 plain path for this embedding");
    }
    if (c->D == 2) {
        const int lim = kMaxLen2d;
        if (g->Mp > lim)
            return fail(FITSNE_E_UNSUPPORTED, "2D grid of " + std::to_string(g->n1d) + " nodes/dim (FFT length " +
                                                  std::to_string(g->Mp) + ") exceeds this build's limit " +
                                                  std::to_string(lim));
    } else {
        const int lim = f32 ? kMaxTot32 / 2 : kMaxTot64;
        if (g->Mp > lim)
            return fail(FITSNE_E_UNSUPPORTED, "1D grid FFT length " + std::to_string(g->Mp) + " exceeds " +
                                                  std::to_string(lim));
    }
    c->last = *g;
    c->last_min = mn;
    c->last_max = mx;
    return FITSNE_OK;
}

// Host-planned geometry of this step into the device record the kernels read (live; kernel spectra
// stale when they must be recomputed).  The cache fields of the captured path are reset, so a later
// captured step recomputes its spectra and twiddles.
static DevGeo host_record(const fitsne_ctx *c, const Geo &g, bool kspec_stale) {
    DevGeo v{};
    v.g = g;
    v.pl = make_plan(g.Mp);
    v.live = 1;
    v.kspec_stale = kspec_stale ? 1 : 0;
    v.mn = c->last_min;
    v.mx = c->last_max;
    return v;
}

static int put_geo(fitsne_ctx *c, const Geo &g, bool kspec_stale) {
    TRY(ensure(c->dgeo, sizeof(DevGeo)));
    launch(c, "set_geo", k_set_geo, dim3(1), dim3(kWave), 0, host_record(c, g, kspec_stale), c->dgeo.as<DevGeo>());
    return FITSNE_OK;
}

// Apply a pending zeroMean to Y in memory (k_center, with the min / max partials a captured step's
// geometry starts from).  Every entry point that reads or replaces Y other than an FP32 update step
// calls this first; the step's readers of Y on the second stream are done by then (the first stream
// joined them before its column pass).
static int materialize(fitsne_ctx *c) {
    if (!c->ctr_lazy) return FITSNE_OK;
    HIPTRY(hipSetDevice(c->device));
    const long n = (long)c->N * c->D;
    const int cblk = part_blocks(n);
    TRY(ensure(c->mmpart, sizeof(double) * 2 * cblk));
    launch(c, "center", k_center<float>, dim3(cblk), dim3(kThreads), 0, c->dgeo.as<const DevGeo>(), c->Y.as<float>(),
           n, c->D, skip_coord(c), c->mean.as<const double>(), c->mmpart.as<double>());
    c->ctr_lazy = false;
    c->mm_parts_valid = true;
    return FITSNE_OK;
}

static int put_step(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    TRY(ensure(c->dstep, sizeof(StepDev)));
    StepDev v{};
    v.mode = mode;
    v.exag = (float)c->exag;
    if (sp) {
        v.momentum = (float)sp->momentum;
        v.lr = (float)sp->learning_rate;
        v.max_step_norm = (float)sp->max_step_norm;
        v.plain_descent = sp->plain_descent;
    }
    launch(c, "set_step", k_set_step, dim3(1), dim3(kWave), 0, v, c->dstep.as<StepDev>());
    return FITSNE_OK;
}

// ============================================================================ kernel dispatch tables

template <int D>
static auto pick_spread_f32(int p) -> decltype(&k_spread_f32<3, D>) {
    switch (p) {
    case 2: return k_spread_f32<2, D>;
    case 3: return k_spread_f32<3, D>;
    case 4: return k_spread_f32<4, D>;
    default: return k_spread_f32<5, D>;
    }
}
template <int D>
static auto pick_spread_f32_v(int p) -> decltype(&k_spread_f32_v<3, D>) {
    switch (p) {
    case 2: return k_spread_f32_v<2, D>;
    case 3: return k_spread_f32_v<3, D>;
    case 4: return k_spread_f32_v<4, D>;
    default: return k_spread_f32_v<5, D>;
    }
}
template <int D>
static auto pick_interp_f32(int p) -> decltype(&k_interp_f32<3, D>) {
    switch (p) {
    case 2: return k_interp_f32<2, D>;
    case 3: return k_interp_f32<3, D>;
    case 4: return k_interp_f32<4, D>;
    default: return k_interp_f32<5, D>;
    }
}
template <int D>
static auto pick_weights(int p) -> decltype(&k_weights_f64<3, D>) {
    switch (p) {
    case 2: return k_weights_f64<2, D>;
    case 3: return k_weights_f64<3, D>;
    case 4: return k_weights_f64<4, D>;
    default: return k_weights_f64<5, D>;
    }
}
template <int D>
static auto pick_spread_f64(int p) -> decltype(&k_spread_f64<3, D>) {
    switch (p) {
    case 2: return k_spread_f64<2, D>;
    case 3: return k_spread_f64<3, D>;
    case 4: return k_spread_f64<4, D>;
    default: return k_spread_f64<5, D>;
    }
}
template <int D>
static auto pick_interp_f64(int p) -> decltype(&k_interp_f64<3, D>) {
    switch (p) {
    case 2: return k_interp_f64<2, D>;
    case 3: return k_interp_f64<3, D>;
    case 4: return k_interp_f64<4, D>;
    default: return k_interp_f64<5, D>;
    }
}

// Row blocks of about equal cost for the windowed attractive kernel (att_rows == kAttBalanced).  One
// workgroup per CU holds a block, all start together, and the kernel lasts as long as its slowest block,
// so the blocks are cut by a cost model instead of a row count: per block, a time per quad of entries,
// per staged window (two barriers and a drain of the quad pipeline) and per small bucket (read from
// global memory), fitted to per-workgroup timestamps at N = 1M (0.00133, 1.5 and 0.4 µs; the staged
// windows explain most of the spread of block times).  The entries of every chunk of kAttChunk rows
// per window come from k_win_count at that block size; a greedy pass over the chunks closes a block
// when adding the next chunk would pass the target cost or 4096 rows, and the smallest target that
// needs no more blocks than the CUs (times the rounds N needs) is found by bisection.  Any partition
// gives the same forces bit for bit (one fixed-point scale for the problem).
constexpr int kAttBalanced = -1;
// Longest row block of the windowed attractive kernel: its LDS holds the block's Y and int64 sums.
static int att_rows_max(const fitsne_ctx *c) { return c->D == 1 ? kAttRows1 : kAttRows; }
constexpr int kAttChunk = 256;

static int balance_blocks(fitsne_ctx *c, int NW, std::vector<int> &rs) {
    const int N = c->N, amax = att_rows_max(c);
    const int nch = (N + kAttChunk - 1) / kAttChunk;
    DBuf cnt, pur, bm;
    TRY(ensure(cnt, sizeof(int) * (size_t)nch * NW));
    TRY(ensure(pur, sizeof(int) * (size_t)nch * NW));
    TRY(ensure(bm, sizeof(float) * (size_t)nch));
    launch(c, "win_count", k_win_count, dim3(nch), dim3(kAttThreads), sizeof(int) * 2 * (size_t)NW,
           c->rowP.as<const int64_t>(), c->colP.as<const uint32_t>(), c->val.as<const float>(), N, kAttChunk, NW,
           (const int *)nullptr, cnt.as<int>(), pur.as<int>(), bm.as<float>());
    std::vector<int> h((size_t)nch * NW);
    HIPTRY(hipMemcpyAsync(h.data(), cnt.p, sizeof(int) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    release(cnt);
    release(pur);
    release(bm);
    // sparse chunks: (window, entries) pairs
    std::vector<int> off(nch + 1, 0), win, ent;
    for (int k = 0; k < nch; k++) {
        for (int w = 0; w < NW; w++) {
            const int e = h[(size_t)k * NW + w];
            if (e > 0) {
                win.push_back(w);
                ent.push_back(e);
            }
        }
        off[k + 1] = (int)win.size();
    }
    constexpr double kQuad = 0.00133, kStaged = 1.5, kSmall = 0.4;
    std::vector<int> acc(NW, 0);
    std::vector<int> touched;
    auto status = [](int e) { const int q = (e + 3) / 4; return q == 0 ? 0 : q < kSmallQuads ? 1 : 2; };
    // greedy partition for target T: block starts into out (if given), returns the number of blocks
    auto greedy = [&](double T, std::vector<int> *out) {
        int nb = 0, rows = 0;
        double cost = 0.0;
        touched.clear();
        if (out) out->assign(1, 0);
        for (int k = 0; k < nch; k++) {
            const int kr = std::min(kAttChunk, N - k * kAttChunk);
            double d = 0.0;
            for (int i = off[k]; i < off[k + 1]; i++) {
                const int s0 = status(acc[win[i]]), s1 = status(acc[win[i]] + ent[i]);
                d += kQuad * ent[i] / 4.0 + (s1 != s0 ? (s1 == 2 ? kStaged : 0.0) - (s0 == 2 ? kStaged : 0.0) +
                                                            (s1 == 1 ? kSmall : 0.0) - (s0 == 1 ? kSmall : 0.0)
                                                      : 0.0);
            }
            if (rows > 0 && (rows + kr > amax || cost + d > T)) {  // close the block before chunk k
                for (int w : touched) acc[w] = 0;
                touched.clear();
                nb++;
                if (out) out->push_back(k * kAttChunk);
                rows = 0;
                cost = 0.0;
                d = 0.0;
                for (int i = off[k]; i < off[k + 1]; i++)
                    d += kQuad * ent[i] / 4.0 + (status(ent[i]) == 2 ? kStaged : status(ent[i]) == 1 ? kSmall : 0.0);
            }
            for (int i = off[k]; i < off[k + 1]; i++) {
                if (acc[win[i]] == 0) touched.push_back(win[i]);
                acc[win[i]] += ent[i];
            }
            rows += kr;
            cost += d;
        }
        for (int w : touched) acc[w] = 0;
        touched.clear();
        if (out) out->push_back(N);
        return nb + 1;
    };
    int cus = 0;
    HIPTRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const long rounds = std::max(1l, ((long)N + (long)cus * amax - 1) / ((long)cus * amax));
    const int want = (int)std::min<long>((long)cus * rounds, (N + kAttChunk - 1) / kAttChunk);
    double lo = 0.0, hi = 1.0;
    while (greedy(hi, nullptr) > want && hi < 1e12) hi *= 2.0;
    for (int it = 0; it < 40; it++) {
        const double mid = 0.5 * (lo + hi);
        if (greedy(mid, nullptr) <= want) hi = mid;
        else lo = mid;
    }
    greedy(hi, &rs);
    return FITSNE_OK;
}

// Bucketed copy of the CSR for k_attr_win: counts per (row block, window), scan, scatter.  Used when
// the entry offsets fit an int and the window cursors fit in LDS; otherwise the CSR kernel runs.
static int build_windows(fitsne_ctx *c) {
    const int N = c->N;
    c->win_ok = false;
    const bool bal = c->att_rows == kAttBalanced;
    const int NW = (N + kWin - 1) / kWin;
    if (c->nnz <= 0 || NW > 8192) return FITSNE_OK;
    std::vector<int> rs;
    if (bal) TRY(balance_blocks(c, NW, rs));
    const int amax = att_rows_max(c);
    const int R = bal ? amax
                      : c->att_rows > 0 ? std::min(amax, c->att_rows)
                                        : std::min(amax, 64 * (int)((((long)N + 255) / 256 + 63) / 64));
    const int NB = bal ? (int)rs.size() - 1 : (N + R - 1) / R;
    const long nbw = (long)NB * NW;
    const long long cap = c->nnz + 3ll * nbw;  // every bucket padded to a multiple of 4 entries
    if (nbw >= (1l << 28) || cap >= (1ll << 31) - 1) return FITSNE_OK;
    if (bal) {
        TRY(ensure(c->wrstart, sizeof(int) * rs.size()));
        HIPTRY(hipMemcpyAsync(c->wrstart.p, rs.data(), sizeof(int) * rs.size(), hipMemcpyHostToDevice, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));  // rs goes out of scope
    }
    const int *rstart = bal ? c->wrstart.as<const int>() : nullptr;
    TRY(ensure(c->wcount, sizeof(int) * nbw));
    TRY(ensure(c->wpure, sizeof(int) * nbw));
    TRY(ensure(c->wstart, sizeof(int) * (nbw + 1)));
    TRY(ensure(c->wbmax, sizeof(float) * (NB + 1)));
    TRY(ensure(c->wkey, sizeof(uint32_t) * (size_t)cap));
    TRY(ensure(c->wval, sizeof(float) * (size_t)cap));
    const size_t lds = sizeof(int) * 2 * (size_t)NW;
    launch(c, "win_count", k_win_count, dim3(NB), dim3(kAttThreads), lds, c->rowP.as<const int64_t>(),
           c->colP.as<const uint32_t>(), c->val.as<const float>(), N, R, NW, rstart, c->wcount.as<int>(),
           c->wpure.as<int>(), c->wbmax.as<float>());
    launch(c, "bmax_all", k_bmax_all, dim3(1), dim3(1024), 0, c->wbmax.as<float>(), NB);
    launch(c, "scan_counts", k_scan_counts, dim3(1), dim3(1024), 0, c->wcount.as<int>(), c->wstart.as<int>(), (int)nbw);
    launch(c, "win_fill", k_win_fill, dim3(NB), dim3(kAttThreads), lds, c->rowP.as<const int64_t>(),
           c->colP.as<const uint32_t>(), c->val.as<const float>(), N, R, NW, rstart, c->wstart.as<const int>(),
           c->wpure.as<const int>(), c->wkey.as<uint32_t>(), c->wval.as<float>());
    HIPTRY(hipGetLastError());
    c->win_R = R;
    c->win_NB = NB;
    c->win_NW = NW;
    c->win_bal = bal;
    c->win_ok = true;
    return FITSNE_OK;
}

static int live_record(fitsne_ctx *c, const DevGeo **out) {
    if (!c->dlive_set) {
        TRY(ensure(c->dlive, sizeof(DevGeo)));
        DevGeo v{};
        v.live = 1;
        launch(c, "set_geo", k_set_geo, dim3(1), dim3(kWave), 0, v, c->dlive.as<DevGeo>());
        c->dlive_set = true;
    }
    *out = c->dlive.as<const DevGeo>();
    return FITSNE_OK;
}

// FP32: the windowed kernel reads the P multiplier from the step record (scale is for the CSR kernel)
// and skips the step when the geometry record dg says so.
// ctr: a pending centring of Y (windowed kernel only).
template <typename TV, int D>
static int launch_attractive(fitsne_ctx *c, TV scale, TV *pos, const DevGeo *dg = nullptr, bool captured = false,
                             const double *ctr = nullptr) {
    const TV inv_df = (TV)(1.0 / c->df);
    const int df1 = c->df == 1.0;
    if constexpr (std::is_same<TV, float>::value) {
        if (!c->win_ok) TRY(build_windows(c));
        if (ctr && !c->win_ok) return fail(FITSNE_E_UNSUPPORTED, "pending centring without the windowed layout");
        if (c->win_ok) {
            const int NB = c->win_NB;
            launch(c, "attractive", k_attr_win<D>, dim3(8 * ((NB + 7) / 8)), dim3(kAttThreads), 0,
                   c->wstart.as<const int>(), c->wpure.as<const int>(), c->wkey.as<const uint32_t>(),
                   c->wval.as<const float>(), c->wbmax.as<const float>(), c->Y.as<const float>(), c->N, c->win_R,
                   c->win_NW, c->win_bal ? c->wrstart.as<const int>() : (const int *)nullptr, NB, inv_df, df1,
                   (float)(0.5 * std::sqrt(c->df)), c->dstep.as<const StepDev>(), captured ? -1.f : (float)scale,
                   dg, pos, ctr);
            return FITSNE_OK;
        }
    }
    const int sw = sw_for(c->nnz, c->N);
    const int rpb = kThreads / sw;
    const int nblk = (c->N + rpb - 1) / rpb;
    dim3 grid(8 * ((nblk + 7) / 8));
    auto rp = c->rowP.as<const int64_t>();
    auto cp = c->colP.as<const uint32_t>();
    auto vp = c->val.as<const TV>();
    auto yp = c->Y.as<const TV>();
    if (sw == 32)
        launch(c, "attractive", k_attractive<TV, int64_t, uint32_t, D, 32>, grid, dim3(kThreads), 0, rp, cp, vp, yp,
               c->N, inv_df, df1, scale, pos);
    else
        launch(c, "attractive", k_attractive<TV, int64_t, uint32_t, D, 16>, grid, dim3(kThreads), 0, rp, cp, vp, yp,
               c->N, inv_df, df1, scale, pos);
    return FITSNE_OK;
}

// ============================================================================ FP32 pipeline

// Largest grids of the captured path (the build's limits): 2D FFT length 2048 (n1d <= 1024), 1D 4096.
constexpr int kCapN1d2 = kMaxLen2d / 2, kCapMp2 = kMaxLen2d;
constexpr int kCapMp1 = kMaxTot32 / 2, kCapN1d1 = kCapMp1 / 2;
// k_kline1d's workgroup: one wave per SIMD, so that it fits in the registers the 1D attractive kernel's
// workgroup leaves on its CU (four waves per SIMD at 88 VGPRs each, of 512).  With two waves per SIMD
// (80 VGPRs each) it did not fit, and waited for that kernel's end.
constexpr int kKline1dThreads = 256;

// Launch sizes of the grid stages: exact for a host-planned step, the build's limits for a captured one.
struct StageDims {
    int Mp;          // FFT length the LDS is sized for
    int gr_rows;     // workgroups of the row passes
    int gr_cols;     // workgroups of the column passes
    int gr_nodes;    // workgroups of the 1D grid conversion
};

// Spread, FFT convolution, Z, interpolation (+ update, zeroMean) of the FP32 form, reading geometry,
// plan and step parameters from the context's device records.  side: the kernel spectra go on the
// second stream (behind the attractive kernel there), so that they run beside the spread and the
// forward row pass; the column pass waits for them.  (Measured at N = 1M:
// this beats computing them on the main stream before or after the spread by ~0.015 ms per step.)
// In 1D host-planned steps the stages up to Z run on the second stream (c->swap_main holds the main
// stream, where the attractive kernel runs), and the interpolation returns to the main stream.
// ctr: the centring still pending on Y (host-planned steps), applied by the spread and the interpolation.
// hv: the host-planned geometry record; the spread and the kernel spectra take it by value and the
// spread stores it for the other passes, so no launch or cross-stream wait stands before them (without
// hv the record is already in device memory: captured steps, or set by put_geo).
template <int D>
static int launch_f32_stages(fitsne_ctx *c, const cx<float> *tw, const StageDims &sd, bool kern, int mode,
                             bool side, const double *ctr = nullptr, const DevGeo *hv = nullptr) {
    const int N = c->N, p = c->p;
    const bool df1 = c->df == 1.0;
    if (hv) TRY(ensure(c->dgeo, sizeof(DevGeo)));  // the spread stores the record there (a fresh context's first step)
    const DevGeo *dg = c->dgeo.as<const DevGeo>();
    if (!dg) return fail(FITSNE_E_ARG, "no geometry record");
    const int kind_q = df1 ? FITSNE_KERNEL_SQUARED_CAUCHY : FITSNE_KERNEL_SQUARED_GENERAL;
    const int kind_1 = FITSNE_KERNEL_GENERAL;
    auto kernel_spectra = [&]() {
        const size_t Mp = (size_t)sd.Mp;
        const size_t lds = sizeof(cx<float>) * (lspan<true>((int)Mp) + Mp);
        if (hv) {
            launch(c, "kern_rows", k_kern_rows_v<float, kKernThreads, kKernTot, true>, dim3(sd.gr_rows),
                   dim3(kKernThreads), lds, *hv, tw, kind_q, kind_1, c->df, c->Trow.as<cx<float>>());
            launch(c, "kern_cols", k_kern_cols_v<float, kKernThreads, kKernTot, true>, dim3(sd.gr_cols),
                   dim3(kKernThreads), lds, *hv, tw, c->Trow.as<const cx<float>>(), 1, c->KQ.as<float>(),
                   c->K1.as<float>(), c->KS.as<float>(), c->KT.as<float>());
        } else {
            launch(c, "kern_rows", k_kern_rows<float, kKernThreads, kKernTot, true>, dim3(sd.gr_rows),
                   dim3(kKernThreads), lds, dg, tw, kind_q, kind_1, c->df, c->Trow.as<cx<float>>());
            launch(c, "kern_cols", k_kern_cols<float, kKernThreads, kKernTot, true>, dim3(sd.gr_cols),
                   dim3(kKernThreads), lds, dg, tw, c->Trow.as<const cx<float>>(), 1, c->KQ.as<float>(),
                   c->K1.as<float>(), c->KS.as<float>(), c->KT.as<float>());
        }
    };
    const bool kside = D == 2 && kern && side && c->overlap;
    if (kside) {
        // the spectra read the record (by value with hv) and the twiddles: wait for the stream only when
        // it wrote either in this step
        if (!hv || c->tw_fresh) {
            HIPTRY(hipEventRecord(c->geo_ev, c->stream));
            HIPTRY(hipStreamWaitEvent(c->astream, c->geo_ev, 0));
        }
        hipStream_t main = c->stream;
        c->stream = c->astream;
        kernel_spectra();
        c->stream = main;
        HIPTRY(hipEventRecord(c->kern_ev, c->astream));
    }
    // 1D: the kernel line (its spectra depend only on the geometry record and the twiddles) on the third
    // stream, after the work queued so far on the grid stages' stream (the twiddle table, the last
    // convolution's read of kline)
    const bool kl_side = D == 1 && hv && side && c->overlap;
    if (kl_side) {
        TRY(ensure(c->kline, sizeof(float) * 3 * (size_t)sd.Mp));
        if (!c->kstream) HIPTRY(hipStreamCreateWithFlags(&c->kstream, hipStreamNonBlocking));
        HIPTRY(hipEventRecord(c->kl_go, c->stream));
        HIPTRY(hipStreamWaitEvent(c->kstream, c->kl_go, 0));
        hipStream_t main = c->stream;
        c->stream = c->kstream;
        launch(c, "kline1d", k_kline1d<float, kKline1dThreads, kCapMp1>, dim3(1), dim3(kKline1dThreads),
               sizeof(cx<float>) * (size_t)sd.Mp, *hv, tw, kind_q, kind_1, c->df, c->kline.as<float>());
        c->stream = main;
        HIPTRY(hipEventRecord(c->kl_ev, c->kstream));
    }
    c->tw_fresh = false;
    const int pts = c->spread_pts;
    if (hv) {
        launch(c, "spread_f32", pick_spread_f32_v<D>(p), dim3((N + pts - 1) / pts), dim3(kSpreadThreads), 0,
               c->Y.as<const float>(), N, pts, *hv, c->dgeo.as<DevGeo>(), c->fixgrid.as<unsigned long long>(), ctr);
    } else {
        launch(c, "spread_f32", pick_spread_f32<D>(p), dim3((N + pts - 1) / pts), dim3(kSpreadThreads), 0,
               c->Y.as<const float>(), N, pts, dg, c->fixgrid.as<unsigned long long>(), ctr);
    }
    if constexpr (D == 2) {
        const size_t Mp = (size_t)sd.Mp;
        if (kern && !kside) kernel_spectra();
        launch(c, "rows_fwd", k_rows_fwd<float, kRowThreads, kRowTot, true>, dim3(sd.gr_rows), dim3(kRowThreads),
               sizeof(cx<float>) * (lspan<true>(2 * (int)Mp) + Mp), dg, tw, (const float *)nullptr, 4, 4, c->spec.as<cx<float>>(),
               c->fixgrid.as<unsigned long long>());
        if (kside) {
            // kern_ev follows the attractive kernel on the side stream, so this one wait also joins it:
            // each cross-stream wait costs about 10 us of queue latency, even on a finished event
            HIPTRY(hipStreamWaitEvent(c->stream, c->kern_ev, 0));
            c->join_pending = false;
        }
        launch(c, "cols_f32", k_cols_f32<kColThreads, kColTot32>, dim3(sd.gr_cols), dim3(kColThreads),
               sizeof(cx<float>) * (lspan<true>(4 * (int)Mp) + Mp), dg, tw, c->spec.as<const cx<float>>(), c->KQ.as<const float>(),
               c->K1.as<const float>(), c->KS.as<const float>(), c->KT.as<const float>(), c->ospec.as<cx<float>>(),
               c->zpart.as<double>());
        // the final Z is computed by the inverse row pass's first workgroup
        launch(c, "rows_inv", k_rows_inv<float, kRowThreads, kRowTot, true>, dim3(sd.gr_rows), dim3(kRowThreads),
               sizeof(cx<float>) * (lspan<true>(2 * (int)Mp) + Mp), dg, tw, c->ospec.as<const cx<float>>(), 3, c->gout.as<float>(), 4,
               c->zpart.as<const double>(), (int)df1, (double)N, c->zbuf.as<double>());
    } else {
        launch(c, "grid_to_f32", k_grid_to_f32, dim3(sd.gr_nodes), dim3(kThreads), 0, dg,
               c->fixgrid.as<unsigned long long>(), c->gin.as<float4>());
        if (kl_side) HIPTRY(hipStreamWaitEvent(c->stream, c->kl_ev, 0));
        launch(c, "conv1d_f32", k_conv1d<float, kConv1dThreads, kMaxTot32>, dim3(1), dim3(kConv1dThreads),
               sizeof(cx<float>) * 2 * (size_t)sd.Mp, dg, tw, kind_q, kind_1, c->df, 1, c->gin.as<const float>(), 4, 3,
               0, c->gout.as<float>(), 4, c->zpart.as<double>(), (const float *)nullptr,
               kl_side ? c->kline.as<const float>() : (const float *)nullptr);
        launch(c, "zfinal", k_zfinal, dim3(1), dim3(kThreads), 0, dg, c->zpart.as<const double>(), (int)df1, (double)N,
               c->zbuf.as<double>());
    }
    const int nblk = (N + kInterpPts - 1) / kInterpPts;
    if (c->swap_main) {
        // back to the main stream, behind the attractive kernel, once the grid stages are done
        HIPTRY(hipEventRecord(c->join_ev, c->stream));
        c->stream = c->swap_main;
        c->swap_main = nullptr;
        HIPTRY(hipStreamWaitEvent(c->stream, c->join_ev, 0));
    }
    if (c->join_pending) {
        HIPTRY(hipStreamWaitEvent(c->stream, c->join_ev, 0));
        c->join_pending = false;
    }
    // host-planned updates: the interpolation also leaves the per-coordinate max / min of the new Y, and
    // mean_mm turns them into those of the centred Y in pinned memory, so that the host plans the next
    // geometry while the centring runs
    const bool mm_host = side && mode == 1;
    launch(c, "interp_f32", pick_interp_f32<D>(p), dim3(nblk), dim3(kThreads), 0, c->Y.as<float>(),
           c->pos.as<const float>(), c->gout.as<const float4>(), dg, c->zbuf.as<const double>(),
           c->dstep.as<const StepDev>(), c->dC.as<float>(), c->uY.as<float>(), c->gains.as<float>(),
           c->part.as<double>(), N, mm_host ? c->mmu.as<double>() : (double *)nullptr, skip_coord(c), ctr);
    if (mode == 1) {
        const long n = (long)N * D;
        const int cblk = part_blocks(n);
        if (mm_host) {
            launch(c, "mean_mm", k_mean_mm<D>, dim3(1), dim3(kThreads), 0, dg, c->part.as<const double>(),
                   c->mmu.as<const double>(), nblk, (double)N, c->Y.as<const float>(), first_coord(c),
                   c->mean.as<double>(), c->h_mm_dev);
            HIPTRY(hipEventRecord(c->mm_event, c->stream));
        } else {
            launch(c, "mean_final", k_mean_final, dim3(1), dim3(kThreads), 0, dg, c->part.as<const double>(), nblk,
                   D, (double)N, c->mean.as<double>());
        }
        if (mm_host && c->win_ok) {
            // zeroMean left pending: the next step's readers apply the mean now in `mean` on load
            c->ctr_lazy = true;
        } else {
            // the min / max partials of the centred Y, for a device-planned step's geometry
            launch(c, "center", k_center<float>, dim3(cblk), dim3(kThreads), 0, dg, c->Y.as<float>(), n, D,
                   skip_coord(c), c->mean.as<const double>(), c->mmpart.as<double>());
            c->ctr_lazy = false;
        }
    }
    return FITSNE_OK;
}

// Buffers of the grid stages for n1d nodes per dimension (FFT length Mp).
template <int D>
static int ensure_f32_buffers(fitsne_ctx *c, int n1d, int Mp) {
    const int N = c->N;
    const int H = Mp / 2 + 1;
    const long nodes = D == 2 ? (long)n1d * n1d : n1d;
    TRY(ensure(c->pos, sizeof(float) * (size_t)N * D));
    TRY(ensure(c->dC, sizeof(float) * (size_t)N * D));
    TRY(ensure(c->part, sizeof(double) * (size_t)blocks_for(N) * D));
    TRY(ensure(c->mean, sizeof(double) * 2));
    TRY(ensure(c->mmpart, sizeof(double) * 2 * part_blocks((long)N * D)));
    TRY(ensure(c->mm, sizeof(double) * 4));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    TRY(ensure(c->mmu, sizeof(double) * 2 * D * (size_t)blocks_for(N)));
    TRY(ensure(c->gin, sizeof(float4) * nodes));
    TRY(ensure(c->gout, sizeof(float4) * nodes));
    if (c->fixgrid.cap < sizeof(unsigned long long) * 4 * (size_t)nodes) {
        // the exact grid is zero between iterations (its reader clears it); fresh memory is zeroed here
        release(c->fixgrid);
        TRY(ensure(c->fixgrid, sizeof(unsigned long long) * 4 * (size_t)nodes));
        const long nw = (long)(c->fixgrid.cap / sizeof(unsigned long long));
        launch(c, "fill", k_fill<unsigned long long>, dim3(std::min(blocks_for(nw), 2048)), dim3(kThreads), 0,
               c->fixgrid.as<unsigned long long>(), nw, 0ull);
    }
    if constexpr (D == 2) {
        const size_t HH = (size_t)H * H;
        const size_t plane = (size_t)H * n1d;
        TRY(ensure(c->Trow, sizeof(cx<float>) * plane));
        TRY(ensure(c->KQ, sizeof(float) * HH));
        TRY(ensure(c->K1, sizeof(float) * HH));
        TRY(ensure(c->KS, sizeof(float) * HH));
        TRY(ensure(c->KT, sizeof(float) * HH));
        TRY(ensure(c->spec, sizeof(cx<float>) * plane * 4));
        TRY(ensure(c->ospec, sizeof(cx<float>) * plane * 3));
        TRY(ensure(c->zpart, sizeof(double) * 4 * H));
    } else {
        TRY(ensure(c->zpart, sizeof(double) * 4));
    }
    return FITSNE_OK;
}

// mode 0: gradient into c->dC (float); mode 1: gradient + optimiser update + zeroMean.  The host plans
// the geometry from the previous step's min / max (fetched while the attractive kernel runs).
template <int D>
static int pipeline_f32(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    TRY(ensure(c->pos, sizeof(float) * (size_t)c->N * D));
    TRY(ensure(c->dstep, sizeof(StepDev)));
    // a pending centring is carried only from one update step to the next (whose min / max the
    // previous one left); otherwise it is applied here, before anything reads Y
    if (c->ctr_lazy && (mode != 1 || !c->mm_pending)) TRY(materialize(c));
    const double *ctr = c->ctr_lazy ? c->mean.as<const double>() : nullptr;
    const DevGeo *live;
    TRY(live_record(c, &live));
    // 2D: the attractive term on the second stream (after everything queued so far), beside the grid
    // stages; the interpolation waits for it.
    // The second stream is made on first use: contexts that only run captured or batched steps never
    // create it, so their streams alone share the device's hardware queues (a second stream per context
    // measured a third fewer batched iterations).
    // In 1D the attractive kernel sets the step's length, so it stays on the main stream, where it starts
    // as soon as the previous step's last kernel ends (a wait across streams costs ~10 us even on a
    // finished event); the grid stages go to the second stream, which waits for the fork, and the
    // interpolation returns to the main stream behind the attractive kernel with one wait for them.
    // The fork's latency also lets the attractive kernel take its CUs before the spread's many small
    // workgroups arrive (when the spread came first, C3 steps took 0.59 against 0.51 ms).
    struct SwapGuard {
        fitsne_ctx *c;
        ~SwapGuard() {
            if (c->swap_main) {
                c->stream = c->swap_main;
                c->swap_main = nullptr;
            }
        }
    } swap_guard{c};
    if (c->overlap && D == 1) {
        if (!c->astream) HIPTRY(hipStreamCreateWithFlags(&c->astream, hipStreamNonBlocking));
        HIPTRY(hipEventRecord(c->fork_ev, c->stream));
        HIPTRY(hipStreamWaitEvent(c->astream, c->fork_ev, 0));
        TRY((launch_attractive<float, D>(c, (float)c->exag, c->pos.as<float>(), live, false, ctr)));
        c->swap_main = c->stream;
        c->stream = c->astream;
    } else if (c->overlap) {
        if (!c->astream) HIPTRY(hipStreamCreateWithFlags(&c->astream, hipStreamNonBlocking));
        HIPTRY(hipEventRecord(c->fork_ev, c->stream));
        HIPTRY(hipStreamWaitEvent(c->astream, c->fork_ev, 0));
        hipStream_t main = c->stream;
        c->stream = c->astream;
        const int rc = launch_attractive<float, D>(c, (float)c->exag, c->pos.as<float>(), live, false, ctr);
        c->stream = main;
        TRY(rc);
        HIPTRY(hipEventRecord(c->join_ev, c->astream));
        c->join_pending = true;
    } else {
        TRY((launch_attractive<float, D>(c, (float)c->exag, c->pos.as<float>(), live, false, ctr)));
    }
    // the step record (read by the interpolation; the attractive kernel took the exaggeration as an
    // argument) is written beside the attractive kernel
    TRY(put_step(c, mode, sp));
    double mn, mx;
    TRY(current_minmax<float>(c, &mn, &mx));
    Geo g;
    TRY(geometry(c, mn, mx, &g));
    const cx<float> *tw;
    TRY(twiddles<float>(c, g.Mp, &tw));
    TRY(ensure_f32_buffers<D>(c, g.n1d, g.Mp));
    const bool kern = D == 2 && !(c->kcache && c->kc_mn == mn && c->kc_mx == mx && c->kc_nb == g.nb);
    const DevGeo hv = host_record(c, g, kern);  // stored by the spread (k_set_geo's role)
    if (kern) {
        c->kcache = true;
        c->kc_mn = mn;
        c->kc_mx = mx;
        c->kc_nb = g.nb;
    }
    const StageDims sd{g.Mp, 8 * ((g.n1d + 7) / 8), 8 * ((g.H + 7) / 8), blocks_for(g.n1d)};
    TRY(launch_f32_stages<D>(c, tw, sd, kern, mode, true, ctr, &hv));
    if (mode == 1) {  // k_mean_mm writes (max, min, Y[first]) of the centred Y to h_mm before mm_event
        c->mm_pending = true;
        c->mm_valid = false;
        c->mm_parts_valid = !c->ctr_lazy;  // k_center's partials
    }
    c->have_grad = true;
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

// One step of the FP32 form with its geometry planned on the device: no host round trip, so the
// sequence can be captured once and replayed (hipGraph).  Grids and buffers are sized for the
// build's limits; a step whose grid exceeds them is refused on the device (DevGeo.live = 0,
// DevGeo.err = 1) and leaves the state unchanged.
template <int D>
static int launch_step_dev(fitsne_ctx *c) {
    const int n1d_max = D == 2 ? kCapN1d2 : kCapN1d1, mp_max = D == 2 ? kCapMp2 : kCapMp1;
    const int cblk = part_blocks((long)c->N * D);
    launch(c, "geometry", k_geometry<float>, dim3(1), dim3(1024), 0, c->mmpart.as<const double>(), cblk,
           c->Y.as<const float>(), first_coord(c), c->permuted ? c->inv_d.as<const int>() : (const int *)nullptr,
           (long)c->N * D, D, c->p, c->ipi, c->mni, n1d_max, mp_max, lagrange_inbox_bound(c->p), c->dgeo.as<DevGeo>());
    launch(c, "twiddles", k_twiddles_dev<float>, dim3(8), dim3(256), 0, c->dgeo.as<const DevGeo>(),
           c->dtw.as<cx<float>>());
    TRY((launch_attractive<float, D>(c, 1.f, c->pos.as<float>(), c->dgeo.as<const DevGeo>(), true)));
    const int H_max = mp_max / 2 + 1;
    const StageDims sd{mp_max, std::min(8 * ((n1d_max + 7) / 8), 512), std::min(8 * ((H_max + 7) / 8), 512),
                       std::min(blocks_for(n1d_max), 64)};
    TRY(launch_f32_stages<D>(c, c->dtw.as<const cx<float>>(), sd, D == 2, 1, false));
    return FITSNE_OK;
}

// ============================================================================ autotuner
// Points per spread workgroup, chosen by timing: each candidate runs the spread on the context's current
// Y and geometry on its own (best of three launches), and the fastest is kept.  The grid, and so every
// result, is the same for every candidate (the spread's waves are the same aligned groups of 128
// points); only the time differs: small problems want more, smaller workgroups, large ones fewer table
// flushes.  A larger candidate is kept only when it is more than 3% faster alone: inside the step the
// spread fills the CUs the attractive kernel frees, which smaller workgroups do better (at N = 2M in 1D,
// 1024 points measured 4% slower per step than 512, though alone the two are within noise).
constexpr int kSpreadPtsCand[] = {512, 1024, 2048, 4096};
constexpr float kSpreadGain = 0.97f;

template <int D>
static int tune_spread(fitsne_ctx *c) {
    const int N = c->N;
    double mn, mx;
    TRY(current_minmax<float>(c, &mn, &mx));
    Geo g;
    TRY(geometry(c, mn, mx, &g));
    TRY(ensure_f32_buffers<D>(c, g.n1d, g.Mp));
    TRY(put_geo(c, g, true));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPTRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(FITSNE_E_HIP, "event creation failed");
    }
    const long nw = (long)(c->fixgrid.cap / sizeof(unsigned long long));
    float best_ms = INFINITY;
    int best = c->spread_pts;
    int rc = FITSNE_OK;
    for (int pts : kSpreadPtsCand) {
        float t = INFINITY;
        for (int rep = 0; rep < 3 && rc == FITSNE_OK; rep++) {
            if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(FITSNE_E_HIP, "event record failed");
            launch(c, "spread_f32", pick_spread_f32<D>(c->p), dim3((N + pts - 1) / pts), dim3(kSpreadThreads), 0,
                   c->Y.as<const float>(), N, pts, c->dgeo.as<const DevGeo>(), c->fixgrid.as<unsigned long long>(),
                   (const double *)nullptr);
            if (hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(FITSNE_E_HIP, "event record failed");
            // the grid is zero between iterations: clear what the trial added
            launch(c, "fill", k_fill<unsigned long long>, dim3(std::min(blocks_for(nw), 2048)), dim3(kThreads), 0,
                   c->fixgrid.as<unsigned long long>(), nw, 0ull);
            float ms = INFINITY;
            if (rc == FITSNE_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(FITSNE_E_HIP, "event timing failed");
            t = std::min(t, ms);
        }
        if (rc != FITSNE_OK) break;
        if (t < kSpreadGain * best_ms) {
            best_ms = t;
            best = pts;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TRY(rc);
    TRY(sync_ctx(c));
    if (best != c->spread_pts) {
        c->spread_pts = best;
        drop_graph(c);
    }
    return FITSNE_OK;
}

// Rows per attractive workgroup, chosen by timing the windowed kernel alone (best of three) for each
// candidate: about N / 256 (enough workgroups for the chip at any N), 1024, 2048 and 4096 rows.  The
// windowed copy of P is rebuilt for each candidate and for the choice.  The forces do not depend on it
// (one fixed-point scale for all rows), only the time: small row blocks re-stage the column windows more
// often, large ones leave CUs idle at small N.
template <int D>
static int tune_attractive(fitsne_ctx *c) {
    if (c->nnz <= 0) return FITSNE_OK;
    const int cand[] = {0, 1024, 2048, 4096, kAttBalanced};
    const DevGeo *live;
    TRY(live_record(c, &live));
    TRY(ensure(c->pos, sizeof(float) * (size_t)c->N * D));
    TRY(ensure(c->dstep, sizeof(StepDev)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPTRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(FITSNE_E_HIP, "event creation failed");
    }
    float best_ms = INFINITY;
    int best = c->att_rows;
    int rc = FITSNE_OK;
    for (int r : cand) {
        c->att_rows = r;
        c->win_ok = false;
        rc = build_windows(c);
        if (rc != FITSNE_OK || !c->win_ok) break;
        float t = INFINITY;
        for (int rep = 0; rep < 3 && rc == FITSNE_OK; rep++) {
            if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(FITSNE_E_HIP, "event record failed");
            if (rc == FITSNE_OK) rc = launch_attractive<float, D>(c, (float)c->exag, c->pos.as<float>(), live);
            if (rc == FITSNE_OK && hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(FITSNE_E_HIP, "event record failed");
            float ms = INFINITY;
            if (rc == FITSNE_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(FITSNE_E_HIP, "event timing failed");
            t = std::min(t, ms);
        }
        if (rc != FITSNE_OK) break;
        if (t < best_ms) {
            best_ms = t;
            best = r;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->att_rows = best;
    c->win_ok = false;
    drop_graph(c);
    TRY(rc);
    TRY(build_windows(c));
    return sync_ctx(c);
}

// ============================================================================ FP64 plain pipeline

// One convolution pass of the plain path: charges `set` (nt terms) spread, convolved with kernel
// spectrum KC (2D) or kernel kind (1D), interpolated into pot[N][nt] (zeroed first).
template <int D>
static int plain_pass(fitsne_ctx *c, const Geo &g, const cx<double> *tw, int set, int nt,
                      const double *KC, int kind1d, int use_k1, DBuf &pot) {
    const int N = c->N;
    const long nodes = D == 2 ? (long)g.n1d * g.n1d : g.n1d;
    TRY(ensure(c->ch, sizeof(double) * (size_t)N * nt));
    TRY(ensure(c->w, sizeof(double) * nodes * nt));
    TRY(ensure(c->V, sizeof(double) * nodes * nt));
    TRY(ensure(pot, sizeof(double) * (size_t)N * nt));
    launch(c, "charges", k_charges<D>, dim3(blocks_for(N)), dim3(kThreads), 0, c->Y.as<const double>(), N, set,
           c->ch.as<double>());
    launch(c, "fill", k_fill<double>, dim3(std::min(blocks_for(nodes * nt), 2048)), dim3(kThreads), 0,
           c->w.as<double>(), nodes * nt, 0.0);
    launch(c, "spread_f64", pick_spread_f64<D>(c->p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->ch.as<const double>(), nt, N, g, c->w.as<double>());
    if constexpr (D == 2) {
        const size_t plane = (size_t)g.H * g.n1d;
        TRY(ensure(c->spec, sizeof(cx<double>) * plane * nt));
        TRY(ensure(c->ospec, sizeof(cx<double>) * plane * nt));
        const int gr_rows = 8 * ((g.n1d + 7) / 8), gr_cols = 8 * ((g.H + 7) / 8);
        const int lines = (nt + 1) / 2;
        launch(c, "rows_fwd64", k_rows_fwd<double, kRowThreads, kRowTot>, dim3(gr_rows), dim3(kRowThreads),
               sizeof(cx<double>) * (lines + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->w.as<const double>(), nt, nt,
               c->spec.as<cx<double>>(), (unsigned long long *)nullptr);
        const int cl = std::min(nt, kColTot64 / g.Mp);
        launch(c, "cols_plain64", k_cols_plain<double, kColThreads64, kColTot64>, dim3(gr_cols), dim3(kColThreads64),
               sizeof(cx<double>) * (cl + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->spec.as<const cx<double>>(), nt, KC,
               c->ospec.as<cx<double>>());
        launch(c, "rows_inv64", k_rows_inv<double, kRowThreads, kRowTot>, dim3(gr_rows), dim3(kRowThreads),
               sizeof(cx<double>) * (lines + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->ospec.as<const cx<double>>(), nt,
               c->V.as<double>(), nt, (const double *)nullptr, 0, 0.0, (double *)nullptr);
    } else {
        launch(c, "conv1d_f64", k_conv1d<double, kConv1dThreads, kMaxTot64>, dim3(1), dim3(kConv1dThreads),
               sizeof(cx<double>) * g.Mp, c->dgeo.as<const DevGeo>(), tw, kind1d, kind1d, c->df, 0, c->w.as<const double>(), nt, nt, use_k1,
               c->V.as<double>(), nt, (double *)nullptr, (const double *)nullptr, (const double *)nullptr);
    }
    launch(c, "fill", k_fill<double>, dim3(std::min(blocks_for((long)N * nt), 2048)), dim3(kThreads), 0,
           pot.as<double>(), (long)N * nt, 0.0);
    launch(c, "interp_f64", pick_interp_f64<D>(c->p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->V.as<const double>(), nt, N, g, pot.as<double>());
    return FITSNE_OK;
}

template <int D>
static int pipeline_f64(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    const int N = c->N, p = c->p;
    const bool df1 = c->df == 1.0;
    TRY(ensure(c->pos, sizeof(double) * (size_t)N * D));
    TRY((launch_attractive<double, D>(c, 1.0, c->pos.as<double>())));
    double mn, mx;
    TRY(current_minmax<double>(c, &mn, &mx));
    Geo g;
    TRY(geometry(c, mn, mx, &g));
    const cx<double> *tw;
    TRY(twiddles<double>(c, g.Mp, &tw));
    TRY(put_geo(c, g, true));

    const double *xs = c->Y.as<const double>(), *ys = nullptr;
    if constexpr (D == 2) {
        TRY(ensure(c->xs, sizeof(double) * (size_t)N));
        TRY(ensure(c->ys, sizeof(double) * (size_t)N));
        launch(c, "split", k_split_f64, dim3(blocks_for(N)), dim3(kThreads), 0, c->Y.as<const double>(), N,
               c->xs.as<double>(), c->ys.as<double>());
        xs = c->xs.as<const double>();
        ys = c->ys.as<const double>();
    }
    TRY(ensure(c->box, sizeof(int) * (size_t)N));
    TRY(ensure(c->wx, sizeof(double) * (size_t)N * p));
    TRY(ensure(c->wy, sizeof(double) * (size_t)N * p));
    launch(c, "weights_f64", pick_weights<D>(p), dim3(blocks_for(N)), dim3(kThreads), 0, xs, ys, N, g,
           (const double *)nullptr, (const double *)nullptr, c->box.as<int>(), c->wx.as<double>(), c->wy.as<double>());

    const int kind_q = df1 ? FITSNE_KERNEL_SQUARED_CAUCHY : FITSNE_KERNEL_SQUARED_GENERAL;
    const double *KQ = nullptr, *K1 = nullptr;
    if constexpr (D == 2) {
        const size_t HH = (size_t)g.H * g.H;
        TRY(ensure(c->Trow, sizeof(cx<double>) * (size_t)g.H * g.n1d));
        TRY(ensure(c->KQ, sizeof(double) * HH));
        TRY(ensure(c->K1, sizeof(double) * HH));
        const int gr_rows = 8 * ((g.n1d + 7) / 8), gr_cols = 8 * ((g.H + 7) / 8);
        launch(c, "kern_rows64", k_kern_rows<double, kKernThreads, kKernTot>, dim3(gr_rows), dim3(kKernThreads),
               sizeof(cx<double>) * (1 + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, kind_q, (int)FITSNE_KERNEL_GENERAL, c->df,
               c->Trow.as<cx<double>>());
        launch(c, "kern_cols64", k_kern_cols<double, kKernThreads, kKernTot>, dim3(gr_cols), dim3(kKernThreads),
               sizeof(cx<double>) * (1 + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->Trow.as<const cx<double>>(), 0, c->KQ.as<double>(),
               c->K1.as<double>(), (double *)nullptr, (double *)nullptr);
        KQ = c->KQ.as<const double>();
        K1 = c->K1.as<const double>();
    }
    int zform, cform;
    DBuf *potF = &c->potA;
    if (df1) {
        TRY((plain_pass<D>(c, g, tw, D == 2 ? 0 : 3, D == 2 ? 4 : 3, KQ, kind_q, 0, c->potA)));
        zform = D == 2 ? 0 : 2;
        cform = D == 2 ? 0 : 2;
        potF = &c->potA;
    } else {
        TRY((plain_pass<D>(c, g, tw, D == 2 ? 1 : 4, D == 2 ? 3 : 2, KQ, kind_q, 0, c->potA)));
        TRY((plain_pass<D>(c, g, tw, 2, 1, K1, FITSNE_KERNEL_GENERAL, 1, c->potB)));
        zform = 1;
        cform = D == 2 ? 1 : 3;
        potF = &c->potA;
    }
    const int nblk = blocks_for(N);
    TRY(ensure(c->part, sizeof(double) * (size_t)nblk * 2));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    launch(c, "zterms64", k_zterms_f64, dim3(nblk), dim3(kThreads), 0, c->Y.as<const double>(),
           (df1 ? c->potA : c->potB).as<const double>(), N, D, zform, c->part.as<double>());
    launch(c, "zsum64", k_sum_final, dim3(1), dim3(kThreads), 0, c->part.as<const double>(), nblk, (double)N,
           c->zbuf.as<double>());
    TRY(ensure(c->dC, sizeof(double) * (size_t)N * D));
    launch(c, "combine64", k_combine_f64, dim3(nblk), dim3(kThreads), 0, c->Y.as<const double>(),
           c->pos.as<const double>(), potF->as<const double>(), c->zbuf.as<const double>(), N, cform,
           c->dC.as<double>());
    if (mode == 1) {
        launch(c, "update64", k_update_f64, dim3(nblk), dim3(kThreads), 0, c->Y.as<double>(), c->uY.as<double>(),
               c->gains.as<double>(), c->dC.as<const double>(), N, D, sp->momentum, sp->learning_rate,
               sp->max_step_norm, sp->plain_descent, c->part.as<double>());
        TRY(ensure(c->mean, sizeof(double) * 2));
        const long n = (long)N * D;
        const int cblk = part_blocks(n);
        TRY(ensure(c->mmpart, sizeof(double) * 2 * cblk));
        TRY(ensure(c->mm, sizeof(double) * 4));
        launch(c, "mean_final", k_mean_final, dim3(1), dim3(kThreads), 0, c->dgeo.as<const DevGeo>(), c->part.as<const double>(), nblk, D,
               (double)N, c->mean.as<double>());
        launch(c, "center64", k_center<double>, dim3(cblk), dim3(kThreads), 0, c->dgeo.as<const DevGeo>(), c->Y.as<double>(), n, D,
               skip_coord(c), c->mean.as<const double>(), c->mmpart.as<double>());
        launch(c, "minmax_final", k_minmax_final<double>, dim3(1), dim3(kWave), 0, c->mmpart.as<const double>(), cblk,
               c->Y.as<const double>(), first_coord(c), c->mm.as<double>());
        HIPTRY(hipMemcpyAsync(c->h_mm, c->mm.p, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipEventRecord(c->mm_event, c->stream));
        c->mm_pending = true;
        c->mm_valid = false;
    }
    c->have_grad = true;
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

// ============================================================================ exact path (theta = 0)

// Fixed grid of the partial sums over the N^2 entries of a dense P.
constexpr int kDenseSumBlocks = 1024;

// Gradient of computeExactGradient (tsne.cpp:1232-1282) at the resident Y over the dense P; mode 1 adds
// the update and zeroMean of the step (tsne.cpp:478-531).  zbuf[0] = Z (the reference's sum_Q).
template <int D>
static int pipeline_exact(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    const int N = c->N;
    const long nd = (long)N * D;
    TRY(ensure(c->att, sizeof(double) * (size_t)nd));
    TRY(ensure(c->rep, sizeof(double) * (size_t)nd));
    TRY(ensure(c->zrow, sizeof(double) * (size_t)N));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    TRY(ensure(c->dC, sizeof(double) * (size_t)nd));
    launch(c, "exact_rows", c->df == 1.0 ? k_exact_rows<D, true, true> : k_exact_rows<D, false, true>,
           dim3((N + kExRows - 1) / kExRows), dim3(kThreads), 0,
           c->val.as<const double>(), c->Y.as<const double>(), N, c->df, c->att.as<double>(), c->rep.as<double>(),
           c->zrow.as<double>());
    launch(c, "zsum64", k_sum_final, dim3(1), dim3(kThreads), 0, c->zrow.as<const double>(), N, 0.0,
           c->zbuf.as<double>());
    const int eblk = std::min(blocks_for(nd), 4096);
    launch(c, "exact_combine", k_exact_combine, dim3(eblk), dim3(kThreads), 0, c->att.as<const double>(),
           c->rep.as<const double>(), c->zbuf.as<const double>(), nd, c->dC.as<double>());
    if (mode == 1) {
        const DevGeo *live;
        TRY(live_record(c, &live));
        const int nblk = blocks_for(N);
        TRY(ensure(c->part, sizeof(double) * (size_t)nblk * D));
        TRY(ensure(c->mean, sizeof(double) * kExMaxDims));
        const int cblk = part_blocks(nd);
        TRY(ensure(c->mmpart, sizeof(double) * 2 * cblk));
        launch(c, "update64", k_update_f64, dim3(nblk), dim3(kThreads), 0, c->Y.as<double>(), c->uY.as<double>(),
               c->gains.as<double>(), c->dC.as<const double>(), N, D, sp->momentum, sp->learning_rate,
               sp->max_step_norm, sp->plain_descent, c->part.as<double>());
        launch(c, "mean_final", k_mean_final, dim3(1), dim3(kThreads), 0, live, c->part.as<const double>(), nblk, D,
               (double)N, c->mean.as<double>());
        launch(c, "center64", k_center<double>, dim3(cblk), dim3(kThreads), 0, live, c->Y.as<double>(), nd, D, -1L,
               c->mean.as<const double>(), c->mmpart.as<double>());
    }
    c->have_grad = true;
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

template <int D>
static int kl_exact_d(fitsne_ctx *c) {
    const int N = c->N;
    const int nblk = (N + kExRows - 1) / kExRows;
    TRY(ensure(c->zrow, sizeof(double) * (size_t)N));
    TRY(ensure(c->klpart, sizeof(double) * (size_t)std::max(nblk, 2)));
    TRY(ensure(c->mean, sizeof(double) * kExMaxDims));
    // sum_Q of the current Y (evaluateError computes its own, tsne.cpp:1298-1308)
    launch(c, "exact_z", c->df == 1.0 ? k_exact_rows<D, true, false> : k_exact_rows<D, false, false>, dim3(nblk), dim3(kThreads), 0, (const double *)nullptr,
           c->Y.as<const double>(), N, c->df, (double *)nullptr, (double *)nullptr, c->zrow.as<double>());
    launch(c, "kl_zsum", k_sum_final, dim3(1), dim3(kThreads), 0, c->zrow.as<const double>(), N, 0.0,
           c->mean.as<double>() + 2);
    launch(c, "exact_kl", c->df == 1.0 ? k_exact_kl<D, true> : k_exact_kl<D, false>, dim3(nblk), dim3(kThreads), 0, c->val.as<const double>(),
           c->Y.as<const double>(), N, c->df, c->mean.as<const double>() + 2, c->klpart.as<double>());
    launch(c, "kl_sum", k_sum_final, dim3(1), dim3(kThreads), 0, c->klpart.as<const double>(), nblk, 0.0,
           c->mean.as<double>());
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

static int kl_exact(fitsne_ctx *c) {
    switch (c->D) {
    case 1: return kl_exact_d<1>(c);
    case 2: return kl_exact_d<2>(c);
    case 3: return kl_exact_d<3>(c);
    default: return kl_exact_d<4>(c);
    }
}

// ============================================================================ Barnes-Hut (nbody_algo = 1)

// Pinned host memory of at least `bytes` for the Barnes-Hut transfers.
static int bh_pinned(fitsne_ctx *c, size_t bytes) {
    if (c->bh_pin_cap >= bytes) return FITSNE_OK;
    TRY(sync_ctx(c));
    if (c->bh_pin) (void)hipHostFree(c->bh_pin);
    c->bh_pin = nullptr;
    c->bh_pin_cap = 0;
    const size_t cap = bytes + bytes / 4;
    if (hipHostMalloc(&c->bh_pin, cap, hipHostMallocDefault) != hipSuccess) {
        c->bh_pin = nullptr;
        return fail(FITSNE_E_ALLOC, "pinned host memory for the Barnes-Hut tree");
    }
    c->bh_pin_cap = cap;
    return FITSNE_OK;
}

template <int D>
static BhTree<D> &bh_tree_for(fitsne_ctx *c) {
    if constexpr (D == 1) return c->bh_tree1;
    else if constexpr (D == 2) return c->bh_tree2;
    else return c->bh_tree3;
}

// The SPTree of the resident Y, built on the host in insertion order (from a cached host copy of Y) and
// uploaded in pre-order, one copy each way through pinned memory.
template <int D>
static int bh_tree_of_Y(fitsne_ctx *c) {
    const int N = c->N;
    const size_t ybytes = sizeof(double) * (size_t)N * D;
    TRY(bh_pinned(c, ybytes));
    HIPTRY(hipMemcpyAsync(c->bh_pin, c->Y.p, ybytes, hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    c->bh_y.resize((size_t)N * D);
    std::memcpy(c->bh_y.data(), c->bh_pin, ybytes);
    BhTree<D> &t = bh_tree_for<D>(c);
    t.build(c->bh_y.data(), N);
    const size_t n = t.nodes.size(), tbytes = bh_flat_bytes(n, D);
    TRY(bh_pinned(c, tbytes));
    bh_flatten(t, BhFlatView(c->bh_pin, n, D));
    TRY(ensure(c->bh_tree_d, tbytes));
    HIPTRY(hipMemcpyAsync(c->bh_tree_d.p, c->bh_pin, tbytes, hipMemcpyHostToDevice, c->stream));
    c->bh_nodes = n;
    return FITSNE_OK;
}

template <int D>
static int bh_walk(fitsne_ctx *c, bool forces) {
    const int N = c->N;
    TRY(ensure(c->bh_Q, sizeof(double) * (size_t)N));
    if (forces) TRY(ensure(c->bh_neg, sizeof(double) * (size_t)N * D));
    const BhFlatView v(c->bh_tree_d.p, c->bh_nodes, D);
    launch(c, "bh_nonedge", k_bh_nonedge<D>, dim3(blocks_for(N)), dim3(kThreads), 0, c->Y.as<const double>(), N,
           (const double *)v.com, (const double *)v.maxw, (const unsigned *)v.cum, (const int *)v.leafpt,
           (const int *)v.isleaf, (const int *)v.skip, c->theta, forces ? c->bh_neg.as<double>() : (double *)nullptr,
           c->bh_Q.as<double>());
    return FITSNE_OK;
}

// computeGradient (tsne.cpp:596-640) at the resident Y; mode 1 adds the update and zeroMean.
template <int D>
static int pipeline_bh(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    const int N = c->N;
    const long nd = (long)N * D;
    TRY(bh_tree_of_Y<D>(c));
    TRY(ensure(c->pos, sizeof(double) * (size_t)nd));
    TRY(ensure(c->dC, sizeof(double) * (size_t)nd));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    launch(c, "bh_edge", k_bh_edge<D>, dim3(blocks_for(N)), dim3(kThreads), 0, c->rowP.as<const int64_t>(),
           c->colP.as<const uint32_t>(), c->val.as<const double>(), c->Y.as<const double>(), N, c->pos.as<double>());
    TRY(bh_walk<D>(c, true));
    launch(c, "bh_sumq", k_seq_sum, dim3(1), dim3(1024), 0, c->bh_Q.as<const double>(), N, c->zbuf.as<double>());
    launch(c, "bh_combine", k_bh_combine, dim3(std::min(blocks_for(nd), 4096)), dim3(kThreads), 0,
           c->pos.as<const double>(), c->bh_neg.as<const double>(), c->zbuf.as<const double>(), nd, c->dC.as<double>());
    if (mode == 1) {
        const DevGeo *live;
        TRY(live_record(c, &live));
        const int nblk = blocks_for(N);
        TRY(ensure(c->part, sizeof(double) * (size_t)nblk * D));
        TRY(ensure(c->mean, sizeof(double) * kExMaxDims));
        const int cblk = part_blocks(nd);
        TRY(ensure(c->mmpart, sizeof(double) * 2 * cblk));
        launch(c, "update64", k_update_f64, dim3(nblk), dim3(kThreads), 0, c->Y.as<double>(), c->uY.as<double>(),
               c->gains.as<double>(), c->dC.as<const double>(), N, D, sp->momentum, sp->learning_rate,
               sp->max_step_norm, sp->plain_descent, c->part.as<double>());
        launch(c, "mean_final", k_mean_final, dim3(1), dim3(kThreads), 0, live, c->part.as<const double>(), nblk, D,
               (double)N, c->mean.as<double>());
        launch(c, "center64", k_center<double>, dim3(cblk), dim3(kThreads), 0, live, c->Y.as<double>(), nd, D, -1L,
               c->mean.as<const double>(), c->mmpart.as<double>());
    }
    c->have_grad = true;
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

// evaluateError, Barnes-Hut form (tsne.cpp:1359-1390): Z of the current Y from the tree (summed over the
// points in index order; the reference keeps one running sum across the points), then the KL over the
// edges with q = 1 / (1 + d^2) / Z.  The result goes to c->mean[0].
template <int D>
static int kl_bh_d(fitsne_ctx *c) {
    const int N = c->N;
    TRY(bh_tree_of_Y<D>(c));
    TRY(bh_walk<D>(c, false));
    TRY(ensure(c->mean, sizeof(double) * kExMaxDims));
    launch(c, "bh_sumq", k_seq_sum, dim3(1), dim3(1024), 0, c->bh_Q.as<const double>(), N, c->mean.as<double>() + 2);
    const long threads = (long)N * 32;
    const int nblk = blocks_for(threads);
    TRY(ensure(c->klpart, sizeof(double) * (size_t)std::max(nblk, 2)));
    launch(c, "kl", k_kl<double, int64_t, uint32_t, D, 32>, dim3(nblk), dim3(kThreads), 0, c->rowP.as<const int64_t>(),
           c->colP.as<const uint32_t>(), c->val.as<const double>(), 1.0, c->Y.as<const double>(), N, 1.0,
           c->mean.as<const double>() + 2, c->klpart.as<double>());
    launch(c, "kl_sum", k_sum_final, dim3(1), dim3(kThreads), 0, c->klpart.as<const double>(), nblk, 0.0,
           c->mean.as<double>());
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

static int run_pipeline(fitsne_ctx *c, int mode, const fitsne_step_params *sp) {
    if (c->N <= 0) return fail(FITSNE_E_ARG, "P and Y must be set before computing a gradient");
    HIPTRY(hipSetDevice(c->device));
    if (c->precision == FITSNE_FP32) return c->D == 2 ? pipeline_f32<2>(c, mode, sp) : pipeline_f32<1>(c, mode, sp);
    if (c->precision == FITSNE_FP64_BH) {
        switch (c->D) {
        case 1: return pipeline_bh<1>(c, mode, sp);
        case 2: return pipeline_bh<2>(c, mode, sp);
        default: return pipeline_bh<3>(c, mode, sp);
        }
    }
    if (c->precision == FITSNE_FP64_EXACT) {
        switch (c->D) {
        case 1: return pipeline_exact<1>(c, mode, sp);
        case 2: return pipeline_exact<2>(c, mode, sp);
        case 3: return pipeline_exact<3>(c, mode, sp);
        default: return pipeline_exact<4>(c, mode, sp);
        }
    }
    return c->D == 2 ? pipeline_f64<2>(c, mode, sp) : pipeline_f64<1>(c, mode, sp);
}

// Z through pinned slot 3: slots 0-2 may hold the (max, min, Y[first]) k_mean_mm left for the next step.
static int read_sumq(fitsne_ctx *c, double *sum_Q) {
    double z[2];
    HIPTRY(hipMemcpyAsync(c->h_mm + 3, c->zbuf.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    z[0] = c->h_mm[3];
    c->last_sumQ = z[0];
    if (sum_Q) *sum_Q = z[0];
    return FITSNE_OK;
}

// ============================================================================ C-ABI: misc

extern "C" const char *fitsne_last_error(void) { return g_err.c_str(); }

extern "C" const char *fitsne_version(void) { return "fitsne-mi355x 0.1.0 (reference 1.2.1)"; }

extern "C" int fitsne_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static int check_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(FITSNE_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(FITSNE_E_ARG, "device index out of range");
    return FITSNE_OK;
}

// ============================================================================ C-ABI: resident tier

extern "C" int fitsne_ctx_create(fitsne_ctx **out, int device, int precision, int no_dims, int n_interpolation_points,
                                 double intervals_per_integer, int min_num_intervals, double df) {
    if (!out) return fail(FITSNE_E_ARG, "ctx out-pointer is null");
    *out = nullptr;
    if (precision != FITSNE_FP32 && precision != FITSNE_FP64_PLAIN && precision != FITSNE_FP64_EXACT &&
        precision != FITSNE_FP64_BH)
        return fail(FITSNE_E_ARG, "unknown precision");
    if (precision == FITSNE_FP64_BH) {
        if (no_dims < 1) return fail(FITSNE_E_ARG, "no_dims must be positive");
        if (no_dims > kBhMaxDims)
            return fail(FITSNE_E_UNSUPPORTED, "the Barnes-Hut path supports no_dims <= 3 in this build");
    } else if (precision == FITSNE_FP64_EXACT) {
        if (no_dims < 1) return fail(FITSNE_E_ARG, "no_dims must be positive");
        if (no_dims > kExMaxDims) return fail(FITSNE_E_UNSUPPORTED, "the exact path supports no_dims <= 4 in this build");
    } else if (no_dims != 1 && no_dims != 2) {
        return fail(FITSNE_E_ARG, "no_dims must be 1 or 2 for the FFT gradient");
    }
    if (n_interpolation_points < 2 || n_interpolation_points > kMaxP)
        return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5] in this build");
    if (!(intervals_per_integer > 0) || min_num_intervals < 1 || !(df > 0))
        return fail(FITSNE_E_ARG, "intervals_per_integer, min_num_intervals and df must be positive");
    TRY(check_device(device));
    HIPTRY(hipSetDevice(device));
    fitsne_ctx *c = new fitsne_ctx();
    c->device = device;
    c->precision = precision;
    c->D = no_dims;
    c->p = n_interpolation_points;
    c->ipi = intervals_per_integer;
    c->mni = min_num_intervals;
    c->df = df;
    // The step's events order work between the context's two streams on one device, or (mm_event) tell
    // the host that k_mean_mm is done, whose pinned-memory result it fences at system scope itself; none
    // needs the system-scope cache flush an event otherwise does when recorded (~5 us in the queue each:
    // the N = 1M step went from 0.300 to 0.294 ms, results bitwise unchanged).
    constexpr unsigned kEvFlags = hipEventDisableTiming | hipEventDisableSystemFence;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork_ev, kEvFlags) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, kEvFlags) != hipSuccess ||
        hipEventCreateWithFlags(&c->geo_ev, kEvFlags) != hipSuccess ||
        hipEventCreateWithFlags(&c->kern_ev, kEvFlags) != hipSuccess ||
        hipEventCreateWithFlags(&c->kl_go, kEvFlags) != hipSuccess ||
        hipEventCreateWithFlags(&c->kl_ev, kEvFlags) != hipSuccess ||
        hipHostMalloc((void **)&c->h_mm, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void **)&c->h_mm_dev, c->h_mm, 0) != hipSuccess ||
        hipEventCreateWithFlags(&c->mm_event, kEvFlags) != hipSuccess) {
        delete c;
        return fail(FITSNE_E_HIP, "stream/event/pinned allocation failed");
    }
    c->own_stream = c->stream;
    *out = c;
    return FITSNE_OK;
}

extern "C" int fitsne_ctx_destroy(fitsne_ctx *c) {
    if (!c) return FITSNE_OK;
    if (c->batch) return fail(FITSNE_E_ARG, "destroy the batch that holds this context first");
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->astream) (void)hipStreamSynchronize(c->astream);
    if (c->kstream) (void)hipStreamSynchronize(c->kstream);
    drop_graph(c);
    DBuf *all[] = {&c->rowP, &c->colP, &c->val, &c->Y, &c->uY, &c->gains, &c->pos, &c->dC, &c->wcount, &c->wstart, &c->wpure,
                   &c->wkey, &c->wval, &c->wbmax, &c->wrstart, &c->knnX, &c->knnXT, &c->knnI, &c->knnD, &c->dgeo,
                   &c->dstep, &c->dtw, &c->dlive, &c->klpart, &c->gin, &c->gout, &c->spec, &c->ospec, &c->Trow, &c->KQ,
                   &c->K1, &c->KS, &c->KT, &c->zpart, &c->zbuf, &c->part, &c->mean, &c->mmpart, &c->mm, &c->mmu, &c->xs,
                   &c->ys, &c->box, &c->wx, &c->wy, &c->ch, &c->w, &c->V, &c->potA, &c->potB,
                   &c->inv_d, &c->colP2, &c->val2, &c->tmp, &c->fixgrid, &c->att, &c->rep, &c->zrow,
                   &c->sy_row, &c->sy_col, &c->sy_val, &c->sy_erow, &c->sy_keys, &c->sy_keys2, &c->sy_vals,
                   &c->sy_vals2, &c->sy_partner, &c->sy_cnt, &c->sy_start, &c->sy_okeys, &c->sy_okeys2, &c->sy_tmp,
                   &c->sy_scol, &c->sy_sval, &c->bh_tree_d, &c->bh_neg, &c->bh_Q, &c->kline};
    for (DBuf *b : all) release(*b);
    for (auto &kv : c->tw32) release(kv.second);
    for (auto &kv : c->tw64) release(kv.second);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto &p : c->pend) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (c->mm_event) (void)hipEventDestroy(c->mm_event);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->geo_ev) (void)hipEventDestroy(c->geo_ev);
    if (c->kern_ev) (void)hipEventDestroy(c->kern_ev);
    if (c->kl_go) (void)hipEventDestroy(c->kl_go);
    if (c->kl_ev) (void)hipEventDestroy(c->kl_ev);
    if (c->astream) (void)hipStreamDestroy(c->astream);
    if (c->kstream) (void)hipStreamDestroy(c->kstream);
    if (c->h_mm) (void)hipHostFree(c->h_mm);
    if (c->bh_pin) (void)hipHostFree(c->bh_pin);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return FITSNE_OK;
}

template <typename TI, typename TC>
static int set_P_impl(fitsne_ctx *c, int N, const TI *row_P, const TC *col_P, const double *val_P) {
    if (!c || !row_P || !col_P || !val_P || N <= 0) return fail(FITSNE_E_ARG, "invalid P arguments");
    if (c->precision == FITSNE_FP64_EXACT)
        return fail(FITSNE_E_ARG, "an exact context takes a dense P (fitsne_set_P_dense / fitsne_exact_affinities)");
    HIPTRY(hipSetDevice(c->device));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    const long long nnz = (long long)row_P[N];
    if (nnz < 0 || (long long)row_P[0] != 0) return fail(FITSNE_E_ARG, "row_P must start at 0");
    for (int i = 0; i < N; i++)
        if ((long long)row_P[i + 1] < (long long)row_P[i]) return fail(FITSNE_E_ARG, "row_P must be non-decreasing");
    std::vector<int64_t> r(N + 1);
    for (int i = 0; i <= N; i++) r[i] = (int64_t)row_P[i];
    std::vector<uint32_t> cc((size_t)nnz);
    for (long long k = 0; k < nnz; k++) {
        const long long j = (long long)col_P[k];
        if (j < 0 || j >= N) return fail(FITSNE_E_ARG, "col_P entry out of range");
        cc[k] = (uint32_t)j;
    }
    TRY(sync_ctx(c));
    TRY(ensure(c->rowP, sizeof(int64_t) * (N + 1)));
    TRY(ensure(c->colP, sizeof(uint32_t) * (size_t)std::max<long long>(nnz, 1)));
    HIPTRY(hipMemcpy(c->rowP.p, r.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
    if (nnz) HIPTRY(hipMemcpy(c->colP.p, cc.data(), sizeof(uint32_t) * nnz, hipMemcpyHostToDevice));
    if (c->precision == FITSNE_FP32) {
        std::vector<float> v((size_t)nnz);
        for (long long k = 0; k < nnz; k++) v[k] = (float)val_P[k];
        TRY(ensure(c->val, sizeof(float) * (size_t)std::max<long long>(nnz, 1)));
        if (nnz) HIPTRY(hipMemcpy(c->val.p, v.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
    } else {
        TRY(ensure(c->val, sizeof(double) * (size_t)std::max<long long>(nnz, 1)));
        if (nnz) HIPTRY(hipMemcpy(c->val.p, val_P, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    if (N != c->N) c->mm_valid = c->mm_pending = false;
    c->N = N;
    c->nnz = nnz;
    c->exag = 1.0;
    c->rowP_host = std::move(r);
    c->perm.resize(N);
    c->inv.resize(N);
    for (int i = 0; i < N; i++) c->perm[i] = c->inv[i] = i;
    c->permuted = false;
    c->win_ok = false;
    drop_graph(c);
    c->mm_parts_valid = false;
    return FITSNE_OK;
}

extern "C" int fitsne_set_P(fitsne_ctx *c, int N, const uint32_t *row_P, const uint32_t *col_P, const double *val_P) {
    return set_P_impl(c, N, row_P, col_P, val_P);
}

extern "C" int fitsne_set_P_i64(fitsne_ctx *c, int N, const int64_t *row_P, const int32_t *col_P,
                                const double *val_P) {
    return set_P_impl(c, N, row_P, col_P, val_P);
}

// The FP64 plain context rescales val_P in place exactly as the reference (`*= coeff`, `/= coeff`,
// tsne.cpp:410-411, :537, :542); the FP32 context keeps the scalar.
__global__ void k_scale_val(double *v, long n, double mul, double div) {
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads) {
        double x = v[e];
        if (div != 1.0) x /= div;
        if (mul != 1.0) x *= mul;
        v[e] = x;
    }
}

extern "C" int fitsne_set_exaggeration(fitsne_ctx *c, double factor) {
    if (!c || !(factor > 0)) return fail(FITSNE_E_ARG, "exaggeration must be positive");
    if (factor == c->exag) return FITSNE_OK;
    HIPTRY(hipSetDevice(c->device));
    if (c->precision != FITSNE_FP32 && c->nnz > 0)
        launch(c, "scale_val", k_scale_val, dim3(std::min(blocks_for(c->nnz), 4096)), dim3(kThreads), 0,
               c->val.as<double>(), (long)c->nnz, factor, c->exag);
    c->exag = factor;
    return FITSNE_OK;
}

// ---------------------------------------------------------------- dense P of the exact path

static int exact_ctx_check(fitsne_ctx *c, int N) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    if (c->precision != FITSNE_FP64_EXACT) return fail(FITSNE_E_ARG, "a dense P needs an FITSNE_FP64_EXACT context");
    if (N < 2) return fail(FITSNE_E_ARG, "N must be at least 2");
    return FITSNE_OK;
}

// State of a context after its P changed (N may differ).
static void dense_P_set(fitsne_ctx *c, int N) {
    if (N != c->N) c->mm_valid = c->mm_pending = false;
    c->N = N;
    c->nnz = (long long)N * N;
    c->exag = 1.0;
    c->have_grad = false;
    c->perm.resize(N);
    c->inv.resize(N);
    for (int i = 0; i < N; i++) c->perm[i] = c->inv[i] = i;
    c->permuted = false;
}

extern "C" int fitsne_set_P_dense(fitsne_ctx *c, int N, const double *P) {
    TRY(exact_ctx_check(c, N));
    if (!P) return fail(FITSNE_E_ARG, "null P");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    const size_t nn = (size_t)N * N;
    TRY(ensure(c->val, sizeof(double) * nn));
    HIPTRY(hipMemcpy(c->val.p, P, sizeof(double) * nn, hipMemcpyHostToDevice));
    dense_P_set(c, N);
    return FITSNE_OK;
}

extern "C" int fitsne_exact_affinities(fitsne_ctx *c, const double *X, int N, int D, double perplexity, double sigma,
                                       int perplexity_list_length, const double *perplexity_list, double *P_out,
                                       double *sum_P) {
    TRY(exact_ctx_check(c, N));
    if (!X || D < 1) return fail(FITSNE_E_ARG, "invalid X");
    std::vector<double> plist;
    if (perplexity > 0) {
        plist.push_back(perplexity);
    } else if (perplexity == 0) {
        if (perplexity_list_length < 1 || !perplexity_list) return fail(FITSNE_E_ARG, "empty perplexity list");
        for (int i = 0; i < perplexity_list_length; i++) {
            if (!(perplexity_list[i] > 0)) return fail(FITSNE_E_ARG, "perplexities must be positive");
            plist.push_back(perplexity_list[i]);
        }
    } else {
        if (!(sigma > 0)) return fail(FITSNE_E_ARG, "a fixed kernel width needs sigma > 0");
        plist.push_back(perplexity);  // the kernel takes the fixed width when perplexity <= 0
    }
    const size_t nd = (size_t)N * D, nn = (size_t)N * N;
    for (size_t i = 0; i < nd; i++)
        if (!std::isfinite(X[i])) return fail(FITSNE_E_ARG, "X contains non-finite values");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    TRY(ensure(c->knnX, sizeof(double) * nd));
    TRY(ensure(c->knnXT, sizeof(double) * nd));
    TRY(ensure(c->w, sizeof(double) * nn));
    TRY(ensure(c->val, sizeof(double) * nn));
    TRY(ensure(c->xs, sizeof(double) * (size_t)N));
    TRY(ensure(c->klpart, sizeof(double) * kDenseSumBlocks));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    HIPTRY(hipMemcpy(c->knnX.p, X, sizeof(double) * nd, hipMemcpyHostToDevice));
    launch(c, "transpose64", k_transpose_f64, dim3((N + 31) / 32, (D + 31) / 32), dim3(256), 0,
           c->knnX.as<const double>(), (long)N, D, c->knnXT.as<double>());
    launch(c, "sqdist64", k_sqdist_f64, dim3(blocks_for(N), (N + kSqRows - 1) / kSqRows), dim3(kThreads), 0,
           c->knnX.as<const double>(), c->knnXT.as<const double>(), N, D, c->w.as<double>());
    // computeGaussianPerplexity's rows (tsne.cpp:1519-1525), the list overload averaging (tsne.cpp:1474-1500)
    const long threads = (long)N * kWave;
    for (size_t i = 0; i < plist.size(); i++)
        launch(c, "d2s_dense", k_d2s, dim3(blocks_for(threads)), dim3(kThreads), 0, c->w.as<const double>(),
               c->val.as<double>(), N, N, plist[i], sigma, 1, (const double *)nullptr, c->xs.as<double>(),
               (int *)nullptr, 1, i > 0 ? 1 : 0);
    const long NN = (long)nn;
    const int sblk = std::min(blocks_for(NN), 4096);
    if (perplexity == 0)
        launch(c, "scale_val", k_scale_val, dim3(sblk), dim3(kThreads), 0, c->val.as<double>(), NN, 1.0,
               (double)plist.size());
    // symmetrisation and normalisation (tsne.cpp:196-216)
    const int T = (N + kSymTile - 1) / kSymTile;
    launch(c, "exact_sym", k_exact_sym, dim3(T, T), dim3(256), 0, c->val.as<double>(), N);
    launch(c, "dense_psum", k_partial_sum, dim3(kDenseSumBlocks), dim3(kThreads), 0, c->val.as<const double>(), NN,
           c->klpart.as<double>());
    launch(c, "dense_sum", k_sum_final, dim3(1), dim3(kThreads), 0, c->klpart.as<const double>(), kDenseSumBlocks,
           0.0, c->zbuf.as<double>());
    launch(c, "dense_div", k_div_by, dim3(sblk), dim3(kThreads), 0, c->val.as<double>(), NN,
           c->zbuf.as<const double>());
    HIPTRY(hipGetLastError());
    double total = 0.0;
    HIPTRY(hipMemcpyAsync(c->h_mm + 3, c->zbuf.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (P_out) HIPTRY(hipMemcpyAsync(P_out, c->val.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    total = c->h_mm[3];
    release(c->w);  // the N x N distance matrix is not needed after the calibration
    dense_P_set(c, N);
    if (sum_P) *sum_P = total;
    return FITSNE_OK;
}

// Spatial order of the points for the FP32 form: a stable counting sort of the internal points by the
// Morton index of their cell on a 512 x 512 (2D) or 2^18 (1D) lattice over the bounding box, so that the
// attractive term's neighbour gathers and the spread touch nearby memory.  Deterministic (host sort).

static inline uint32_t morton2(uint32_t x, uint32_t y) {
    uint32_t k = 0;
    for (int b = 0; b < 16; b++) k |= ((x >> b) & 1u) << (2 * b) | ((y >> b) & 1u) << (2 * b + 1);
    return k;
}

static int reorder_points(fitsne_ctx *c) {
    const int N = c->N, D = c->D;
    if (c->precision != FITSNE_FP32 || N < 2) return FITSNE_OK;
    TRY(sync_ctx(c));
    drop_graph(c);
    c->mm_parts_valid = false;
    std::vector<float> y((size_t)N * D);
    HIPTRY(hipMemcpy(y.data(), c->Y.p, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
    float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int i = 0; i < N; i++)
        for (int d = 0; d < D; d++) {
            lo[d] = std::min(lo[d], y[(size_t)i * D + d]);
            hi[d] = std::max(hi[d], y[(size_t)i * D + d]);
        }
    const int G = D == 2 ? 512 : 1 << 18;
    const int nkeys = D == 2 ? G * G : G;
    std::vector<int> key(N), cnt(nkeys + 1, 0);
    for (int i = 0; i < N; i++) {
        uint32_t q[2] = {0, 0};
        for (int d = 0; d < D; d++) {
            const float span = hi[d] - lo[d];
            const float t = span > 0 ? (y[(size_t)i * D + d] - lo[d]) / span : 0.f;
            int v = (int)(t * G);
            q[d] = (uint32_t)std::min(std::max(v, 0), G - 1);
        }
        key[i] = D == 2 ? (int)morton2(q[0], q[1]) : (int)q[0];
        cnt[key[i] + 1]++;
    }
    for (int k = 0; k < nkeys; k++) cnt[k + 1] += cnt[k];
    std::vector<int> rel(N), irel(N);
    for (int i = 0; i < N; i++) rel[cnt[key[i]]++] = i;  // stable
    for (int r = 0; r < N; r++) irel[rel[r]] = r;
    std::vector<int64_t> rown(N + 1);
    rown[0] = 0;
    for (int r = 0; r < N; r++) rown[r + 1] = rown[r] + (c->rowP_host[rel[r] + 1] - c->rowP_host[rel[r]]);
    // device: permute the CSR (rows and column numbering) and the per-point state
    TRY(ensure(c->tmp, sizeof(int) * 2 * (size_t)N));
    TRY(ensure(c->colP2, sizeof(uint32_t) * (size_t)std::max<long long>(c->nnz, 1)));
    TRY(ensure(c->val2, sizeof(float) * (size_t)std::max<long long>(c->nnz, 1)));
    DBuf rowN;
    TRY(ensure(rowN, sizeof(int64_t) * (N + 1)));
    int *d_rel = c->tmp.as<int>(), *d_irel = c->tmp.as<int>() + N;
    HIPTRY(hipMemcpy(d_rel, rel.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(d_irel, irel.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(rowN.p, rown.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
    launch(c, "permute_csr", k_permute_csr<float>, dim3(blocks_for((long)N * kWave)), dim3(kThreads), 0,
           c->rowP.as<const int64_t>(), c->colP.as<const uint32_t>(), c->val.as<const float>(),
           rowN.as<const int64_t>(), (const int *)d_rel, (const int *)d_irel, N, c->colP2.as<uint32_t>(),
           c->val2.as<float>());
    DBuf *state[3] = {&c->Y, &c->uY, &c->gains};
    DBuf scratch;
    TRY(ensure(scratch, sizeof(float) * (size_t)N * D));
    for (DBuf *b : state) {
        launch(c, "gather_rows", k_gather_rows<float>, dim3(blocks_for((long)N * D)), dim3(kThreads), 0,
               b->as<const float>(), (const int *)d_rel, N, D, scratch.as<float>());
        std::swap(*b, scratch);
    }
    TRY(sync_ctx(c));
    release(scratch);
    std::swap(c->colP, c->colP2);
    std::swap(c->val, c->val2);
    std::swap(c->rowP, rowN);
    release(rowN);
    release(c->colP2);
    release(c->val2);
    // compose with the previous order
    std::vector<int> perm(N);
    for (int r = 0; r < N; r++) perm[r] = c->perm[rel[r]];
    c->perm.swap(perm);
    for (int r = 0; r < N; r++) c->inv[c->perm[r]] = r;
    c->rowP_host.swap(rown);
    TRY(ensure(c->inv_d, sizeof(int) * (size_t)N));
    HIPTRY(hipMemcpy(c->inv_d.p, c->inv.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    c->permuted = true;
    c->win_ok = false;
    c->mm_valid = c->mm_pending = false;
    c->kcache = false;
    return FITSNE_OK;
}


extern "C" int fitsne_reorder(fitsne_ctx *c) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    if (c->N <= 0 || !c->Y.p) return fail(FITSNE_E_ARG, "set P and Y before reordering");
    HIPTRY(hipSetDevice(c->device));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    return reorder_points(c);
}

extern "C" int fitsne_set_theta(fitsne_ctx *c, double theta) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    if (c->precision != FITSNE_FP64_BH) return fail(FITSNE_E_ARG, "theta is a parameter of the Barnes-Hut context");
    if (!(theta > 0)) return fail(FITSNE_E_ARG, "theta must be positive (theta = 0 is the exact context)");
    c->theta = theta;
    return FITSNE_OK;
}

extern "C" int fitsne_set_reorder(fitsne_ctx *c, int enable) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    c->reorder_on_set_Y = enable != 0;
    return FITSNE_OK;
}

// threshold below which set_Y keeps the caller's order.  Small embeddings are sorted too: the spread's
// runs of same-box points (and so its atomics) depend on the order whatever the size (measured: 64
// unsorted problems of 2,000 to 65,000 points made the batched spread five times slower per point).
constexpr int kReorderMinN = 2048;

extern "C" int fitsne_set_Y(fitsne_ctx *c, const double *Y) {
    if (!c || !Y) return fail(FITSNE_E_ARG, "null argument");
    if (c->N <= 0) return fail(FITSNE_E_ARG, "set P before Y");
    const size_t n = (size_t)c->N * c->D;
    for (size_t e = 0; e < n; e++)
        if (!std::isfinite(Y[e])) return fail(FITSNE_E_ARG, "Y contains non-finite values");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    c->ctr_lazy = false;  // Y is replaced: a pending centring is void
    const int D = c->D;
    if (c->precision == FITSNE_FP32) {
        std::vector<float> y(n), z(n, 0.f), one(n, 1.f);
        for (size_t r = 0; r < (size_t)c->N; r++)
            for (int d = 0; d < D; d++) y[r * D + d] = (float)Y[(size_t)c->perm[r] * D + d];
        TRY(ensure(c->Y, sizeof(float) * n));
        TRY(ensure(c->uY, sizeof(float) * n));
        TRY(ensure(c->gains, sizeof(float) * n));
        HIPTRY(hipMemcpy(c->Y.p, y.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(c->uY.p, z.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(c->gains.p, one.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    } else {
        std::vector<double> z(n, 0.0), one(n, 1.0);
        TRY(ensure(c->Y, sizeof(double) * n));
        TRY(ensure(c->uY, sizeof(double) * n));
        TRY(ensure(c->gains, sizeof(double) * n));
        HIPTRY(hipMemcpy(c->Y.p, Y, sizeof(double) * n, hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(c->uY.p, z.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(c->gains.p, one.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    c->mm_valid = c->mm_pending = false;
    c->have_grad = false;
    c->kcache = false;
    drop_graph(c);
    c->mm_parts_valid = false;
    if (c->precision == FITSNE_FP32 && c->reorder_on_set_Y && c->N >= kReorderMinN) TRY(reorder_points(c));
    return FITSNE_OK;
}

extern "C" int fitsne_get_Y(fitsne_ctx *c, double *Y) {
    if (!c || !Y) return fail(FITSNE_E_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    TRY(check_graph(c));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    const size_t n = (size_t)c->N * c->D;
    if (c->precision == FITSNE_FP32) {
        std::vector<float> y(n);
        HIPTRY(hipMemcpy(y.data(), c->Y.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        const int D = c->D;
        for (size_t r = 0; r < (size_t)c->N; r++)
            for (int d = 0; d < D; d++) Y[(size_t)c->perm[r] * D + d] = y[r * D + d];
    } else {
        HIPTRY(hipMemcpy(Y, c->Y.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return FITSNE_OK;
}

extern "C" int fitsne_gradient(fitsne_ctx *c, double *dC, double *sum_Q) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    TRY(check_graph(c));
    TRY(run_pipeline(c, 0, nullptr));
    TRY(read_sumq(c, sum_Q));
    if (dC) {
        const size_t n = (size_t)c->N * c->D;
        if (c->precision == FITSNE_FP32) {
            std::vector<float> d(n);
            HIPTRY(hipMemcpy(d.data(), c->dC.p, sizeof(float) * n, hipMemcpyDeviceToHost));
            const int D = c->D;
            for (size_t r = 0; r < (size_t)c->N; r++)
                for (int e = 0; e < D; e++) dC[(size_t)c->perm[r] * D + e] = d[r * D + e];
        } else {
            HIPTRY(hipMemcpy(dC, c->dC.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        }
    }
    return FITSNE_OK;
}

// ============================================================================ captured steps

static void drop_graph(fitsne_ctx *c) {
    c->layout_gen++;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// Capture one device-planned step of the FP32 form (launch_step_dev) on the context's stream.  All
// buffers are allocated first at the build's largest grid; nothing is allocated during capture.
static int capture_step(fitsne_ctx *c) {
    drop_graph(c);
    c->kcache = false;
    if (c->precision != FITSNE_FP32) return fail(FITSNE_E_UNSUPPORTED, "captured steps are for the FP32 form");
    if (c->N <= 0) return fail(FITSNE_E_ARG, "P and Y must be set before a step");
    HIPTRY(hipSetDevice(c->device));
    const int D = c->D;
    TRY(prepare_captured(c));
    TRY(sync_ctx(c));
    const bool timing = c->timing;
    c->timing = false;
    HIPTRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = D == 2 ? launch_step_dev<2>(c) : launch_step_dev<1>(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = timing;
    if (rc != FITSNE_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != FITSNE_OK) return rc;
        HIPTRY(e);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        HIPTRY(ie);
    }
    c->graph = g;
    c->gexec = x;
    c->graph_gen++;
    return FITSNE_OK;
}

// Min / max partials of the resident Y (what a captured step's geometry starts from).
static int ensure_partials(fitsne_ctx *c) {
    if (c->mm_parts_valid) return FITSNE_OK;
    const long n = (long)c->N * c->D;
    const int nblk = part_blocks(n);
    TRY(ensure(c->mmpart, sizeof(double) * 2 * nblk));
    launch(c, "minmax_partial", k_minmax_partial<float>, dim3(nblk), dim3(kThreads), 0, c->Y.as<const float>(), n,
           skip_coord(c), c->mmpart.as<double>());
    c->mm_parts_valid = true;
    return FITSNE_OK;
}

// After captured steps: the host's min / max must be recomputed, and a refused step is an error.
// The captured step rewrites the kernel-spectrum buffers, so the host-planned step's cache of them is void.
static void after_graph_steps(fitsne_ctx *c) {
    c->graph_ran = true;
    c->have_grad = true;
    c->mm_valid = false;
    c->mm_pending = false;
    c->kcache = false;
}

static int check_graph(fitsne_ctx *c) {
    if (!c->graph_ran) return FITSNE_OK;
    DevGeo v;
    HIPTRY(hipMemcpyAsync(&v, c->dgeo.p, sizeof(DevGeo), hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    c->graph_ran = false;
    c->last = v.g;
    c->last_min = v.mn;
    c->last_max = v.mx;
    if (v.err == 2) {
        launch(c, "clear_err", k_clear_err, dim3(1), dim3(kWave), 0, c->dgeo.as<DevGeo>());
        return fail(FITSNE_E_UNSUPPORTED, "a captured step's first point lay so far below the grid (the 2D min rule "
                                          "leaves it out) that the spread had no usable fixed-point scale; the "
                                          "embedding was left at the previous step");
    }
    if (v.err) {
        launch(c, "clear_err", k_clear_err, dim3(1), dim3(kWave), 0, c->dgeo.as<DevGeo>());
        return fail(FITSNE_E_UNSUPPORTED, "a captured step's grid exceeded this build's limits (FFT length " +
                                              std::to_string(c->D == 2 ? kCapMp2 : kCapMp1) +
                                              "); the embedding was left at the previous step");
    }
    return FITSNE_OK;
}

extern "C" int fitsne_autotune(fitsne_ctx *c, int *spread_points) {
    if (!c) return fail(FITSNE_E_ARG, "null argument");
    if (c->precision != FITSNE_FP32) return fail(FITSNE_E_ARG, "the autotuner is for FP32 contexts");
    if (c->N <= 0 || !c->Y.p) return fail(FITSNE_E_ARG, "set P and Y before tuning");
    HIPTRY(hipSetDevice(c->device));
    TRY(check_graph(c));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    TRY(c->D == 2 ? tune_spread<2>(c) : tune_spread<1>(c));
    TRY(c->D == 2 ? tune_attractive<2>(c) : tune_attractive<1>(c));
    if (spread_points) *spread_points = c->spread_pts;
    return FITSNE_OK;
}

extern "C" int fitsne_autotune_stats(fitsne_ctx *c, int *spread_points, int *attractive_rows) {
    if (!c) return fail(FITSNE_E_ARG, "null argument");
    if (spread_points) *spread_points = c->spread_pts;
    if (attractive_rows) *attractive_rows = c->win_ok ? (c->win_bal ? kAttBalanced : c->win_R) : c->att_rows;
    return FITSNE_OK;
}

extern "C" int fitsne_set_attractive_rows(fitsne_ctx *c, int rows) {
    if (!c) return fail(FITSNE_E_ARG, "null argument");
    if (rows != 0 && rows != kAttBalanced && (rows < kWave || rows > att_rows_max(c) || rows % kWave != 0))
        return fail(FITSNE_E_ARG, "attractive rows per workgroup must be 0 (automatic), -1 (balanced blocks) or a "
                                  "multiple of 64 in [64, 4096] (8192 in 1D)");
    HIPTRY(hipSetDevice(c->device));
    TRY(materialize(c));  // the next step may lay P out without the windows the pending centring needs
    TRY(sync_ctx(c));
    if (rows != c->att_rows) {
        c->att_rows = rows;
        c->win_ok = false;
        drop_graph(c);
    }
    return FITSNE_OK;
}

extern "C" int fitsne_set_spread_points(fitsne_ctx *c, int points) {
    if (!c) return fail(FITSNE_E_ARG, "null argument");
    if (points < 2 * kSpreadThreads || points > 8192 || points % (2 * kSpreadThreads) != 0)
        return fail(FITSNE_E_ARG, "spread points per workgroup must be a multiple of 512 in [512, 8192]");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    if (points != c->spread_pts) {
        c->spread_pts = points;
        drop_graph(c);
    }
    return FITSNE_OK;
}

extern "C" int fitsne_set_overlap(fitsne_ctx *c, int enable) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    c->overlap = enable != 0;
    return FITSNE_OK;
}

extern "C" int fitsne_set_graph(fitsne_ctx *c, int enable) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    if (enable && c->precision != FITSNE_FP32) return fail(FITSNE_E_UNSUPPORTED, "captured steps are for the FP32 form");
    // a captured step may still run: wait for it (and report a refused one) before its graph goes
    HIPTRY(hipSetDevice(c->device));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    TRY(check_graph(c));
    drop_graph(c);
    c->use_graph = enable != 0;
    return FITSNE_OK;
}

extern "C" int fitsne_step(fitsne_ctx *c, const fitsne_step_params *prm, double *sum_Q) {
    if (!c || !prm) return fail(FITSNE_E_ARG, "null argument");
    if (c->use_graph) {
        if (c->N <= 0) return fail(FITSNE_E_ARG, "P and Y must be set before a step");
        HIPTRY(hipSetDevice(c->device));
        TRY(materialize(c));
        if (!c->gexec) TRY(capture_step(c));
        TRY(ensure_partials(c));
        TRY(put_step(c, 1, prm));
        HIPTRY(hipGraphLaunch(c->gexec, c->stream));
        after_graph_steps(c);
        if (sum_Q) {
            TRY(check_graph(c));
            TRY(read_sumq(c, sum_Q));
        }
        return FITSNE_OK;
    }
    TRY(run_pipeline(c, 1, prm));
    if (sum_Q) TRY(read_sumq(c, sum_Q));
    return FITSNE_OK;
}

// ============================================================================ batched embeddings

// A batch steps many FP32 contexts together.  When they are all 2D with the same interpolation settings
// and df (the usual seed or perplexity sweep), the batch is fused: one captured graph whose every stage is
// one launch over all problems (fitsne_batch.hpp), on the batch's stream, which the members use as their
// own while they belong to it.  Otherwise each member's captured step is launched on its own stream.
// The step parameters of each iteration go to the device through a ring of pinned slots: a slot is
// refilled only once the copy issued from it kStepSlots iterations earlier has completed.
constexpr int kStepSlots = 4;

struct fitsne_batch {
    std::vector<fitsne_ctx *> ctx;
    bool fused = false;
    int device = 0, p = 3, mni = 50;
    double df = 1.0, ipi = 1.0;
    hipStream_t stream = nullptr, astream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    hipEvent_t slot_ev[kStepSlots] = {};
    bool slot_used[kStepSlots] = {};
    StepDev *h_steps = nullptr;  // pinned, kStepSlots x n
    int slot = 0;
    DBuf dprob, dsteps, dsched, dpref, dspre;
    int tot_attr = 0, tot_spread = 0, tot_interp = 0, tot_center = 0;
    int g_kern[kBsClasses] = {}, g_rows[kBsClasses] = {}, g_cols[kBsClasses] = {};
    std::vector<long> gens;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

static void batch_drop_graph(fitsne_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// Buffers of a device-planned step at the build's largest 2D grid, and a fresh geometry record (no
// kernel spectra or twiddles cached): what a captured step of this context, alone or in a batch, needs.
static int prepare_captured(fitsne_ctx *c) {
    if (!c->win_ok) TRY(build_windows(c));
    if (!c->win_ok) return fail(FITSNE_E_UNSUPPORTED, "captured steps need the windowed attractive layout");
    const int D = c->D;
    TRY(ensure(c->dgeo, sizeof(DevGeo)));
    TRY(ensure(c->dstep, sizeof(StepDev)));
    TRY(ensure(c->dtw, sizeof(cx<float>) * (size_t)(D == 2 ? kCapMp2 : kCapMp1)));
    TRY(D == 2 ? ensure_f32_buffers<2>(c, kCapN1d2, kCapMp2) : ensure_f32_buffers<1>(c, kCapN1d1, kCapMp1));
    DevGeo v{};
    v.live = 1;
    launch(c, "set_geo", k_set_geo, dim3(1), dim3(kWave), 0, v, c->dgeo.as<DevGeo>());
    return FITSNE_OK;
}

template <typename K>
static int persistent_blocks(K kernel, int threads, size_t lds, int cus, int *out) {
    int per = 0;
    HIPTRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(kernel), threads, lds));
    *out = std::max(1, per) * cus;
    return FITSNE_OK;
}

// LDS of the batched FFT passes for FFT lengths up to Mp.
struct BatchLds {
    size_t kern, rows, cols;
};
static BatchLds batch_lds(int Mp) {
    return {sizeof(cx<float>) * (lspan<true>(Mp) + (size_t)Mp), sizeof(cx<float>) * (lspan<true>(2 * Mp) + (size_t)Mp),
            sizeof(cx<float>) * (lspan<true>(4 * Mp) + (size_t)Mp)};
}
// Length class CLS of the batched FFT passes: the buffer of each pass in FFT-length units (the short
// lines fill it with several lines per item) and the workgroup size of each pass.
template <int CLS>
struct BatchClass {
    static constexpr int MP = CLS == 0 ? kBatchItemMp : kCapMp2;       // kernel-spectrum and row passes
    static constexpr int MPC = CLS == 0 ? kBatchSmallMp : kCapMp2;     // column pass
    static constexpr int KT = kKernThreads, RT = kRowThreads, CT = kColThreads;
};
static_assert(kColThreads == kColZLanes, "the column pass's Z lanes are the single-context workgroup");

// The five FFT passes of one length class (CLS), each one persistent launch.
template <int CLS>
static void batch_fft_class(fitsne_batch *b, hipStream_t st, const BatchProb *bp, BatchSched *sch, int *pref,
                            int pass, int kind_q, int kind_1) {
    using C = BatchClass<CLS>;
    const int n = (int)b->ctx.size();
    const BatchLds L = batch_lds(C::MP);
    switch (pass) {
    case kBsKernRows:
        kb_kern_rows<C::KT, C::MP, CLS><<<dim3(b->g_kern[CLS]), dim3(C::KT), L.kern, st>>>(bp, n, sch, pref, kind_q,
                                                                                          kind_1, b->df);
        break;
    case kBsKernCols:
        kb_kern_cols<C::KT, C::MP, CLS><<<dim3(b->g_kern[CLS]), dim3(C::KT), L.kern, st>>>(bp, n, sch, pref);
        break;
    case kBsRowsFwd:
        kb_rows_fwd<C::RT, 2 * C::MP, CLS><<<dim3(b->g_rows[CLS]), dim3(C::RT), L.rows, st>>>(bp, n, sch, pref);
        break;
    case kBsCols:
        kb_cols<C::CT, 4 * C::MPC, CLS><<<dim3(b->g_cols[CLS]), dim3(C::CT), batch_lds(C::MPC).cols, st>>>(bp, n, sch,
                                                                                                      pref);
        break;
    default:
        kb_rows_inv<C::RT, 2 * C::MP, CLS><<<dim3(b->g_rows[CLS]), dim3(C::RT), L.rows, st>>>(bp, n, sch, pref);
        break;
    }
}

template <int CLS>
static int batch_class_blocks(fitsne_batch *b, int cus) {
    using C = BatchClass<CLS>;
    const BatchLds L = batch_lds(C::MP);
    TRY(persistent_blocks(kb_kern_rows<C::KT, C::MP, CLS>, C::KT, L.kern, cus, &b->g_kern[CLS]));
    TRY(persistent_blocks(kb_rows_fwd<C::RT, 2 * C::MP, CLS>, C::RT, L.rows, cus, &b->g_rows[CLS]));
    TRY(persistent_blocks(kb_cols<C::CT, 4 * C::MPC, CLS>, C::CT, batch_lds(C::MPC).cols, cus, &b->g_cols[CLS]));
    return FITSNE_OK;
}

template <int P>
static int batch_capture_p(fitsne_batch *b, const BatchProb *bp, BatchSched *sch, int *pref, const int *spre) {
    const int n = (int)b->ctx.size();
    const bool df1 = b->df == 1.0;
    const int kind_q = df1 ? FITSNE_KERNEL_SQUARED_CAUCHY : FITSNE_KERNEL_SQUARED_GENERAL;
    const int kind_1 = FITSNE_KERNEL_GENERAL;
    const int *apre = spre, *sppre = spre + (n + 1), *ipre = spre + 2 * (n + 1), *cpre = spre + 3 * (n + 1);
    hipStream_t st = b->stream;
    kb_geometry<<<dim3(n), dim3(1024), 0, st>>>(bp, 2, P, b->ipi, b->mni, kCapN1d2, kCapMp2, lagrange_inbox_bound(P));
    // the schedule first: a workgroup launched behind the attractive term waits for one of its
    // workgroups (160 KB of LDS each) to leave a CU
    kb_sched<<<dim3(1), dim3(1024), 0, st>>>(bp, n, sch, pref);
    HIPTRY(hipEventRecord(b->fork_ev, st));
    HIPTRY(hipStreamWaitEvent(b->astream, b->fork_ev, 0));
    kb_attr<<<dim3(b->tot_attr), dim3(kAttThreads), 0, b->astream>>>(bp, n, apre, (float)(1.0 / b->df), (int)df1,
                                                                      (float)(0.5 * std::sqrt(b->df)));
    HIPTRY(hipEventRecord(b->join_ev, b->astream));
    auto passes = [&](int pass) {
        batch_fft_class<0>(b, st, bp, sch, pref, pass, kind_q, kind_1);
        batch_fft_class<1>(b, st, bp, sch, pref, pass, kind_q, kind_1);
    };
    passes(kBsKernRows);
    passes(kBsKernCols);
    kb_spread<P><<<dim3(b->tot_spread), dim3(kSpreadThreads), 0, st>>>(bp, n, sppre);
    passes(kBsRowsFwd);
    passes(kBsCols);
    passes(kBsRowsInv);
    kb_zfinal<<<dim3(n), dim3(kThreads), 0, st>>>(bp, (int)df1);
    HIPTRY(hipStreamWaitEvent(st, b->join_ev, 0));
    kb_interp<P><<<dim3(b->tot_interp), dim3(kThreads), 0, st>>>(bp, n, ipre);
    kb_mean<<<dim3(n), dim3(kThreads), 0, st>>>(bp);
    kb_center<<<dim3(b->tot_center), dim3(kThreads), 0, st>>>(bp, n, cpre);
    return FITSNE_OK;
}

// (Re)build the fused batch: every member's buffers and records, the problem table, the workgroup
// prefixes of the fixed-size stages, the persistent passes' sizes, and the captured graph.
static int batch_prepare(fitsne_batch *b) {
    batch_drop_graph(b);
    const int n = (int)b->ctx.size();
    HIPTRY(hipSetDevice(b->device));
    HIPTRY(hipStreamSynchronize(b->stream));
    for (fitsne_ctx *c : b->ctx) {
        if (c->N <= 0 || !c->Y.p) return fail(FITSNE_E_ARG, "P and Y must be set before a batched step");
        // the largest row blocks: the batch fills the chip with problems, so each workgroup should
        // re-stage the column windows as rarely as possible (the forces do not depend on it)
        if (c->att_rows != kAttRows) {
            c->att_rows = kAttRows;
            c->win_ok = false;
            drop_graph(c);
        }
        TRY(prepare_captured(c));
    }
    TRY(ensure(b->dsteps, sizeof(StepDev) * n));
    std::vector<BatchProb> bp(n);
    std::vector<int> spre(4 * (size_t)(n + 1), 0);
    int *apre = spre.data(), *sppre = apre + (n + 1), *ipre = apre + 2 * (n + 1), *cpre = apre + 3 * (n + 1);
    for (int i = 0; i < n; i++) {
        fitsne_ctx *c = b->ctx[i];
        BatchProb &q = bp[i];
        q.dg = c->dgeo.as<DevGeo>();
        q.sd = b->dsteps.as<const StepDev>() + i;
        q.Y = c->Y.as<float>();
        q.pos = c->pos.as<float>();
        q.dC = c->dC.as<float>();
        q.uY = c->uY.as<float>();
        q.gains = c->gains.as<float>();
        q.N = c->N;
        q.spread_pts = c->spread_pts;
        q.cblk = part_blocks((long)c->N * 2);
        q.first = first_coord(c);
        q.inv = c->permuted ? c->inv_d.as<const int>() : nullptr;
        q.wstart = c->wstart.as<const int>();
        q.wpure = c->wpure.as<const int>();
        q.wkey = c->wkey.as<const uint32_t>();
        q.wval = c->wval.as<const float>();
        q.wbmax = c->wbmax.as<const float>();
        q.win_R = c->win_R;
        q.win_NW = c->win_NW;
        q.fixgrid = c->fixgrid.as<unsigned long long>();
        q.tw = c->dtw.as<cx<float>>();
        q.Trow = c->Trow.as<cx<float>>();
        q.spec = c->spec.as<cx<float>>();
        q.ospec = c->ospec.as<cx<float>>();
        q.KQ = c->KQ.as<float>();
        q.K1 = c->K1.as<float>();
        q.KS = c->KS.as<float>();
        q.KT = c->KT.as<float>();
        q.gout = c->gout.as<float4>();
        q.zpart = c->zpart.as<double>();
        q.zbuf = c->zbuf.as<double>();
        q.part = c->part.as<double>();
        q.mean = c->mean.as<double>();
        q.mmpart = c->mmpart.as<double>();
        const long a = (c->N + c->win_R - 1) / c->win_R, sp = (c->N + c->spread_pts - 1) / c->spread_pts;
        const long ip = (c->N + kInterpPts - 1) / kInterpPts, ce = q.cblk;
        if ((long)apre[i] + a > INT32_MAX / 2 || (long)sppre[i] + sp > INT32_MAX / 2 ||
            (long)ipre[i] + ip > INT32_MAX / 2)
            return fail(FITSNE_E_UNSUPPORTED, "batch too large for one launch per stage");
        apre[i + 1] = apre[i] + (int)a;
        sppre[i + 1] = sppre[i] + (int)sp;
        ipre[i + 1] = ipre[i] + (int)ip;
        cpre[i + 1] = cpre[i] + (int)ce;
    }
    b->tot_attr = apre[n];
    b->tot_spread = sppre[n];
    b->tot_interp = ipre[n];
    b->tot_center = cpre[n];
    TRY(ensure(b->dprob, sizeof(BatchProb) * n));
    TRY(ensure(b->dspre, sizeof(int) * spre.size()));
    TRY(ensure(b->dsched, sizeof(BatchSched)));
    TRY(ensure(b->dpref, sizeof(int) * kBsStages * kBsClasses * (size_t)(n + 1)));
    HIPTRY(hipMemcpyAsync(b->dprob.p, bp.data(), sizeof(BatchProb) * n, hipMemcpyHostToDevice, b->stream));
    HIPTRY(hipMemcpyAsync(b->dspre.p, spre.data(), sizeof(int) * spre.size(), hipMemcpyHostToDevice, b->stream));
    HIPTRY(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
    int cus = 0;
    HIPTRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    TRY(batch_class_blocks<0>(b, cus));
    TRY(batch_class_blocks<1>(b, cus));
    const BatchProb *dbp = b->dprob.as<const BatchProb>();
    auto *sch = b->dsched.as<BatchSched>();
    int *pref = b->dpref.as<int>();
    const int *dspre = b->dspre.as<const int>();
    HIPTRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc;
    switch (b->p) {
    case 2: rc = batch_capture_p<2>(b, dbp, sch, pref, dspre); break;
    case 3: rc = batch_capture_p<3>(b, dbp, sch, pref, dspre); break;
    case 4: rc = batch_capture_p<4>(b, dbp, sch, pref, dspre); break;
    default: rc = batch_capture_p<5>(b, dbp, sch, pref, dspre); break;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != FITSNE_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != FITSNE_OK) return rc;
        HIPTRY(e);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        HIPTRY(ie);
    }
    b->graph = g;
    b->gexec = x;
    b->gens.resize(n);
    for (int i = 0; i < n; i++) b->gens[i] = b->ctx[i]->layout_gen;
    return FITSNE_OK;
}

static void batch_release(fitsne_batch *b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->astream) (void)hipStreamSynchronize(b->astream);
    for (fitsne_ctx *c : b->ctx)
        if (c->batch == b) {
            c->stream = c->own_stream;
            c->batch = nullptr;
        }
    batch_drop_graph(b);
    release(b->dprob);
    release(b->dsteps);
    release(b->dsched);
    release(b->dpref);
    release(b->dspre);
    for (auto &e : b->slot_ev)
        if (e) (void)hipEventDestroy(e);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    if (b->join_ev) (void)hipEventDestroy(b->join_ev);
    if (b->h_steps) (void)hipHostFree(b->h_steps);
    if (b->astream) (void)hipStreamDestroy(b->astream);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

extern "C" int fitsne_batch_create(fitsne_batch **out, fitsne_ctx **ctxs, int n) {
    if (!out || !ctxs || n <= 0) return fail(FITSNE_E_ARG, "invalid batch");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || ctxs[i]->precision != FITSNE_FP32 || ctxs[i]->device != ctxs[0]->device)
            return fail(FITSNE_E_ARG, "a batch holds FP32 contexts on one device");
        if (ctxs[i]->batch) return fail(FITSNE_E_ARG, "a context belongs to another batch");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return fail(FITSNE_E_ARG, "a context appears twice in the batch");
    }
    // every problem gets its own spread workgroup size (results do not depend on it); the attractive row
    // blocks of a fused batch are set below
    for (int i = 0; i < n; i++) {
        fitsne_ctx *c = ctxs[i];
        if (c->N <= 0 || !c->Y.p) return fail(FITSNE_E_ARG, "set P and Y before batching");
        HIPTRY(hipSetDevice(c->device));
        TRY(check_graph(c));
        TRY(materialize(c));
        TRY(sync_ctx(c));
        TRY(c->D == 2 ? tune_spread<2>(c) : tune_spread<1>(c));
    }
    auto *b = new fitsne_batch();
    b->ctx.assign(ctxs, ctxs + n);
    const fitsne_ctx *c0 = ctxs[0];
    b->device = c0->device;
    b->p = c0->p;
    b->df = c0->df;
    b->ipi = c0->ipi;
    b->mni = c0->mni;
    bool fused = true;
    for (int i = 0; i < n; i++) {
        const fitsne_ctx *c = ctxs[i];
        fused = fused && c->D == 2 && c->p == b->p && c->df == b->df && c->ipi == b->ipi && c->mni == b->mni &&
                c->N > 0 && c->nnz > 0;
    }
    if (!fused) {
        *out = b;
        return FITSNE_OK;
    }
    b->fused = true;
    int rc = FITSNE_OK;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->astream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
        hipHostMalloc((void **)&b->h_steps, sizeof(StepDev) * kStepSlots * (size_t)n, hipHostMallocDefault) != hipSuccess)
        rc = fail(FITSNE_E_HIP, "batch stream/event/pinned allocation failed");
    for (int s = 0; s < kStepSlots && rc == FITSNE_OK; s++)
        if (hipEventCreateWithFlags(&b->slot_ev[s], hipEventDisableTiming) != hipSuccess)
            rc = fail(FITSNE_E_HIP, "batch event creation failed");
    if (rc == FITSNE_OK) {
        // the members use the batch's stream from now on, so their own calls stay ordered with its steps
        for (fitsne_ctx *c : b->ctx) {
            if (sync_ctx(c) != FITSNE_OK) rc = FITSNE_E_HIP;
            c->stream = b->stream;
            c->batch = b;
        }
    }
    if (rc == FITSNE_OK) rc = batch_prepare(b);
    if (rc != FITSNE_OK) {
        const std::string msg = g_err;
        batch_release(b);
        delete b;
        g_err = msg;
        return rc;
    }
    *out = b;
    return FITSNE_OK;
}

extern "C" int fitsne_batch_destroy(fitsne_batch *b) {
    if (!b) return FITSNE_OK;
    if (b->fused) {
        (void)hipSetDevice(b->device);
        batch_release(b);
    }
    delete b;
    return FITSNE_OK;
}

extern "C" int fitsne_batch_step(fitsne_batch *b, const fitsne_step_params *prm) {
    if (!b || !prm) return fail(FITSNE_E_ARG, "null argument");
    const int n = (int)b->ctx.size();
    if (!b->fused) {
        for (int i = 0; i < n; i++) {
            fitsne_ctx *c = b->ctx[i];
            HIPTRY(hipSetDevice(c->device));
            c->use_graph = true;
            TRY(materialize(c));
            if (!c->gexec) TRY(capture_step(c));
            TRY(ensure_partials(c));
            TRY(put_step(c, 1, &prm[i]));
            HIPTRY(hipGraphLaunch(c->gexec, c->stream));
            after_graph_steps(c);
        }
        return FITSNE_OK;
    }
    HIPTRY(hipSetDevice(b->device));
    bool stale = !b->gexec;
    for (int i = 0; i < n && !stale; i++) stale = b->ctx[i]->layout_gen != b->gens[i];
    if (stale) TRY(batch_prepare(b));
    for (fitsne_ctx *c : b->ctx) {
        TRY(materialize(c));
        TRY(ensure_partials(c));
    }
    const int s = b->slot;
    if (b->slot_used[s]) HIPTRY(hipEventSynchronize(b->slot_ev[s]));
    StepDev *h = b->h_steps + (size_t)s * n;
    for (int i = 0; i < n; i++) {
        StepDev v{};
        v.mode = 1;
        v.exag = (float)b->ctx[i]->exag;
        v.momentum = (float)prm[i].momentum;
        v.lr = (float)prm[i].learning_rate;
        v.max_step_norm = (float)prm[i].max_step_norm;
        v.plain_descent = prm[i].plain_descent;
        h[i] = v;
    }
    HIPTRY(hipMemcpyAsync(b->dsteps.p, h, sizeof(StepDev) * n, hipMemcpyHostToDevice, b->stream));
    HIPTRY(hipEventRecord(b->slot_ev[s], b->stream));
    b->slot_used[s] = true;
    b->slot = (s + 1) % kStepSlots;
    HIPTRY(hipGraphLaunch(b->gexec, b->stream));
    for (fitsne_ctx *c : b->ctx) after_graph_steps(c);
    return FITSNE_OK;
}

extern "C" int fitsne_kl(fitsne_ctx *c, double *C) {
    if (!c || !C) return fail(FITSNE_E_ARG, "null argument");
    TRY(check_graph(c));
    TRY(materialize(c));  // before the KL's sum overwrites the mean
    if (c->precision == FITSNE_FP64_EXACT || c->precision == FITSNE_FP64_BH) {
        if (c->N <= 0) return fail(FITSNE_E_ARG, "P and Y must be set before the KL divergence");
        HIPTRY(hipSetDevice(c->device));
        if (c->precision == FITSNE_FP64_EXACT) TRY(kl_exact(c));
        else TRY(c->D == 1 ? kl_bh_d<1>(c) : c->D == 2 ? kl_bh_d<2>(c) : kl_bh_d<3>(c));
        HIPTRY(hipMemcpyAsync(c->h_mm + 3, c->mean.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TRY(sync_ctx(c));
        *C = c->h_mm[3];
        return FITSNE_OK;
    }
    if (!c->have_grad) return fail(FITSNE_E_ARG, "KL needs the Z of a previous gradient");
    HIPTRY(hipSetDevice(c->device));
    const int sw = sw_for(c->nnz, c->N);
    (void)sw;
    const long threads = (long)c->N * 32;
    const int nblk = blocks_for(threads);
    TRY(ensure(c->klpart, sizeof(double) * (size_t)std::max(nblk, 2)));
    TRY(ensure(c->mean, sizeof(double) * 2));
    const double pscale = c->precision == FITSNE_FP32 ? c->exag : 1.0;
    auto rp = c->rowP.as<const int64_t>();
    auto cp = c->colP.as<const uint32_t>();
    if (c->precision == FITSNE_FP32) {
        if (c->D == 2)
            launch(c, "kl", k_kl<float, int64_t, uint32_t, 2, 32>, dim3(nblk), dim3(kThreads), 0, rp, cp,
                   c->val.as<const float>(), pscale, c->Y.as<const float>(), c->N, c->df, c->zbuf.as<const double>(),
                   c->klpart.as<double>());
        else
            launch(c, "kl", k_kl<float, int64_t, uint32_t, 1, 32>, dim3(nblk), dim3(kThreads), 0, rp, cp,
                   c->val.as<const float>(), pscale, c->Y.as<const float>(), c->N, c->df, c->zbuf.as<const double>(),
                   c->klpart.as<double>());
    } else {
        if (c->D == 2)
            launch(c, "kl", k_kl<double, int64_t, uint32_t, 2, 32>, dim3(nblk), dim3(kThreads), 0, rp, cp,
                   c->val.as<const double>(), pscale, c->Y.as<const double>(), c->N, c->df,
                   c->zbuf.as<const double>(), c->klpart.as<double>());
        else
            launch(c, "kl", k_kl<double, int64_t, uint32_t, 1, 32>, dim3(nblk), dim3(kThreads), 0, rp, cp,
                   c->val.as<const double>(), pscale, c->Y.as<const double>(), c->N, c->df,
                   c->zbuf.as<const double>(), c->klpart.as<double>());
    }
    launch(c, "kl_sum", k_sum_final, dim3(1), dim3(kThreads), 0, c->klpart.as<const double>(), nblk, 0.0,
           c->mean.as<double>());
    HIPTRY(hipMemcpyAsync(c->h_mm + 3, c->mean.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    *C = c->h_mm[3];
    return FITSNE_OK;
}

extern "C" int fitsne_sync(fitsne_ctx *c) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    HIPTRY(hipSetDevice(c->device));
    TRY(sync_ctx(c));
    return check_graph(c);
}

extern "C" int fitsne_set_timing(fitsne_ctx *c, int enable) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    TRY(sync_ctx(c));
    c->timing = enable != 0;
    c->stats.clear();
    return FITSNE_OK;
}

extern "C" int fitsne_kernel_stats(fitsne_ctx *c, int cap, char *names, int *counts, double *total_ms) {
    if (!c) return fail(FITSNE_E_ARG, "null context");
    TRY(sync_ctx(c));
    int i = 0;
    for (auto &kv : c->stats) {
        if (i >= cap) break;
        if (names) {
            std::strncpy(names + 48 * i, kv.first.c_str(), 47);
            names[48 * i + 47] = 0;
        }
        if (counts) counts[i] = kv.second.count;
        if (total_ms) total_ms[i] = kv.second.ms;
        i++;
    }
    return i;
}

extern "C" int fitsne_last_geometry(fitsne_ctx *c, double *out) {
    if (!c || !out) return fail(FITSNE_E_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    TRY(check_graph(c));
    out[0] = c->last.nb;
    out[1] = c->last.n1d;
    out[2] = c->last.Mp;
    out[3] = c->last_min;
    out[4] = c->last_max;
    out[5] = (double)c->nnz;
    return FITSNE_OK;
}

extern "C" long long fitsne_debug_read(fitsne_ctx *c, const char *name, void *dst, long long bytes) {
    if (!c || !name || !dst || bytes < 0) return fail(FITSNE_E_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    TRY(materialize(c));
    TRY(sync_ctx(c));
    const std::string n(name);
    DBuf *b = n == "grid_in" ? &c->gin : n == "grid_out" ? &c->gout : n == "kq" ? &c->KQ : n == "k1" ? &c->K1
            : n == "ks" ? &c->KS : n == "kt" ? &c->KT : n == "zpart" ? &c->zpart : n == "zbuf" ? &c->zbuf
            : n == "pos" ? &c->pos : n == "dC" ? &c->dC : nullptr;
    if (!b) return fail(FITSNE_E_ARG, "unknown buffer name");
    const long long nb = std::min<long long>(bytes, (long long)b->cap);
    if (nb > 0 && b->p) HIPTRY(hipMemcpy(dst, b->p, (size_t)nb, hipMemcpyDeviceToHost));
    return nb;
}

// ============================================================================ C-ABI: mirror tier

static std::mutex g_mirror_mu;
static fitsne_ctx *g_mirror = nullptr;

static int mirror_ctx(fitsne_ctx **out, int D, int p, double ipi, int mni, double df) {
    if (!g_mirror) TRY(fitsne_ctx_create(&g_mirror, 0, FITSNE_FP64_PLAIN, D, p, ipi, mni, df));
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    if (!(ipi > 0) || mni < 1 || !(df > 0)) return fail(FITSNE_E_ARG, "invalid FFT settings");
    g_mirror->D = D;
    g_mirror->p = p;
    g_mirror->ipi = ipi;
    g_mirror->mni = mni;
    g_mirror->df = df;
    g_mirror->kcache = false;
    *out = g_mirror;
    return FITSNE_OK;
}

template <typename TI, typename TC>
static int mirror_gradient(const TI *row_P, const TC *col_P, const double *val_P, const double *Y, int N, int D,
                           double *dC, int p, double ipi, int mni, double df, double *sum_Q) {
    if (!row_P || !col_P || !val_P || !Y || !dC || N <= 0) return fail(FITSNE_E_ARG, "invalid arguments");
    if (D != 1 && D != 2) return fail(FITSNE_E_UNSUPPORTED, "the FFT gradient supports 1 or 2 dimensions");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, D, p, ipi, mni, df));
    TRY(set_P_impl(c, N, row_P, col_P, val_P));
    TRY(fitsne_set_Y(c, Y));
    double z;
    TRY(fitsne_gradient(c, dC, &z));
    if (sum_Q) *sum_Q = z;
    return FITSNE_OK;
}

extern "C" int fitsne_compute_fft_gradient(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                                           const double *Y, int N, int D, double *dC, int n_interpolation_points,
                                           double intervals_per_integer, int min_num_intervals, unsigned nthreads,
                                           double df, double *sum_Q) {
    (void)nthreads;
    return mirror_gradient(row_P, col_P, val_P, Y, N, D, dC, n_interpolation_points, intervals_per_integer,
                           min_num_intervals, df, sum_Q);
}

extern "C" int fitsne_compute_fft_gradient_i64(const int64_t *row_P, const int32_t *col_P, const double *val_P,
                                               const double *Y, int N, int D, double *dC, int n_interpolation_points,
                                               double intervals_per_integer, int min_num_intervals,
                                               unsigned nthreads, double df, double *sum_Q) {
    (void)nthreads;
    return mirror_gradient(row_P, col_P, val_P, Y, N, D, dC, n_interpolation_points, intervals_per_integer,
                           min_num_intervals, df, sum_Q);
}

// Geometry of a precompute/n_body_fft pair from the caller's box layout.
static Geo mirror_geo(int D, int p, int nb, double mn, double bw0, const double *lower, const double *upper) {
    Geo g = make_geo(D, p, mn, mn + bw0 * nb, nb);
    g.bw0 = bw0;
    g.h = (1 / (double)p) * bw0;
    if (lower && upper) {
        g.mn = lower[0];
        g.bw = upper[0] - lower[0];
    }
    return g;
}

static int kernel_kind_ok(int kernel) {
    return kernel == FITSNE_KERNEL_SQUARED_CAUCHY || kernel == FITSNE_KERNEL_GENERAL ||
           kernel == FITSNE_KERNEL_SQUARED_GENERAL;
}

extern "C" int fitsne_precompute_2d(double x_max, double x_min, double y_max, double y_min, int n_boxes,
                                    int n_interpolation_points, int kernel, double *box_lower_bounds,
                                    double *box_upper_bounds, double *y_tilde_spacings, double *y_tilde,
                                    double *x_tilde, double *fft_kernel_tilde, double df) {
    (void)y_max;
    if (n_boxes < 1 || !kernel_kind_ok(kernel) || !box_lower_bounds || !box_upper_bounds || !y_tilde_spacings ||
        !y_tilde || !x_tilde || !fft_kernel_tilde)
        return fail(FITSNE_E_ARG, "invalid precompute_2d arguments");
    const int p = n_interpolation_points;
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    // box bounds, spacings, nodes: the reference's loops (nbodyfft.cpp:15-46)
    const int ntb = n_boxes * n_boxes;
    const double box_width = (x_max - x_min) / (double)n_boxes;
    for (int i = 0; i < n_boxes; i++)
        for (int j = 0; j < n_boxes; j++) {
            box_lower_bounds[i * n_boxes + j] = j * box_width + x_min;
            box_upper_bounds[i * n_boxes + j] = (j + 1) * box_width + x_min;
            box_lower_bounds[ntb + i * n_boxes + j] = i * box_width + y_min;
            box_upper_bounds[ntb + i * n_boxes + j] = (i + 1) * box_width + y_min;
        }
    double h = 1 / (double)p;
    y_tilde_spacings[0] = h / 2;
    for (int i = 1; i < p; i++) y_tilde_spacings[i] = y_tilde_spacings[i - 1] + h;
    const int n1d = p * n_boxes;
    h = h * box_width;
    x_tilde[0] = x_min + h / 2;
    y_tilde[0] = y_min + h / 2;
    for (int i = 1; i < n1d; i++) {
        x_tilde[i] = x_tilde[i - 1] + h;
        y_tilde[i] = y_tilde[i - 1] + h;
    }
    // kernel spectrum on the GPU: C(m, k) for m, k in [0, H), FP64, written into fft_kernel_tilde
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 2, p, 1.0, 1, df));
    HIPTRY(hipSetDevice(c->device));
    Geo g = mirror_geo(2, p, n_boxes, x_min, box_width, nullptr, nullptr);
    if (g.Mp > kMaxLen2d) return fail(FITSNE_E_UNSUPPORTED, "grid too large for this build");
    const cx<double> *tw;
    TRY(twiddles<double>(c, g.Mp, &tw));
    TRY(put_geo(c, g, true));
    const size_t HH = (size_t)g.H * g.H;
    TRY(ensure(c->Trow, sizeof(cx<double>) * (size_t)g.H * g.n1d));
    TRY(ensure(c->KQ, sizeof(double) * HH));
    TRY(ensure(c->K1, sizeof(double) * HH));
    const int gr_rows = 8 * ((g.n1d + 7) / 8), gr_cols = 8 * ((g.H + 7) / 8);
    launch(c, "kern_rows64", k_kern_rows<double, kKernThreads, kKernTot>, dim3(gr_rows), dim3(kKernThreads),
           sizeof(cx<double>) * (1 + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, kernel, kernel, df, c->Trow.as<cx<double>>());
    launch(c, "kern_cols64", k_kern_cols<double, kKernThreads, kKernTot>, dim3(gr_cols), dim3(kKernThreads),
           sizeof(cx<double>) * (1 + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->Trow.as<const cx<double>>(), 0, c->KQ.as<double>(),
           c->K1.as<double>(), (double *)nullptr, (double *)nullptr);
    HIPTRY(hipMemcpyAsync(fft_kernel_tilde, c->KQ.p, sizeof(double) * HH, hipMemcpyDeviceToHost, c->stream));
    return sync_ctx(c);
}

extern "C" int fitsne_n_body_fft_2d(int N, int n_terms, const double *xs, const double *ys, const double *chargesQij,
                                    int n_boxes, int n_interpolation_points, const double *box_lower_bounds,
                                    const double *box_upper_bounds, const double *y_tilde_spacings,
                                    const double *fft_kernel_tilde, double *potentialQij, unsigned nthreads) {
    (void)nthreads;
    if (N <= 0 || n_terms < 1 || n_boxes < 1 || !xs || !ys || !chargesQij || !box_lower_bounds ||
        !box_upper_bounds || !y_tilde_spacings || !fft_kernel_tilde || !potentialQij)
        return fail(FITSNE_E_ARG, "invalid n_body_fft_2d arguments");
    if (n_terms > 4) return fail(FITSNE_E_UNSUPPORTED, "n_terms must be in [1, 4] in this build");
    const int p = n_interpolation_points;
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 2, p, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    const int ntb = n_boxes * n_boxes;
    Geo g = mirror_geo(2, p, n_boxes, box_lower_bounds[0], box_upper_bounds[0] - box_lower_bounds[0],
                       box_lower_bounds, box_upper_bounds);
    for (int i = 0; i < p; i++) g.sp[i] = y_tilde_spacings[i];
    for (int i = 0; i < p; i++) {
        g.den[i] = 1;
        for (int j = 0; j < p; j++)
            if (i != j) g.den[i] *= g.sp[i] - g.sp[j];
    }
    if (g.Mp > kMaxLen2d) return fail(FITSNE_E_UNSUPPORTED, "grid too large for this build");
    const cx<double> *tw;
    TRY(twiddles<double>(c, g.Mp, &tw));
    TRY(put_geo(c, g, true));
    // caller's per-axis lower bounds: x lower of column j at [j], y lower of row i at [ntb + i*nb]
    std::vector<double> lo(2 * (size_t)n_boxes);
    for (int j = 0; j < n_boxes; j++) lo[j] = box_lower_bounds[j];
    for (int i = 0; i < n_boxes; i++) lo[n_boxes + i] = box_lower_bounds[ntb + i * n_boxes];
    c->N = N;
    const size_t HH = (size_t)g.H * g.H;
    TRY(ensure(c->xs, sizeof(double) * N));
    TRY(ensure(c->ys, sizeof(double) * N));
    TRY(ensure(c->mean, sizeof(double) * 2 * n_boxes));
    TRY(ensure(c->ch, sizeof(double) * (size_t)N * n_terms));
    TRY(ensure(c->potA, sizeof(double) * (size_t)N * n_terms));
    TRY(ensure(c->KQ, sizeof(double) * HH));
    HIPTRY(hipMemcpy(c->xs.p, xs, sizeof(double) * N, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->ys.p, ys, sizeof(double) * N, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->mean.p, lo.data(), sizeof(double) * 2 * n_boxes, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->ch.p, chargesQij, sizeof(double) * (size_t)N * n_terms, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->potA.p, potentialQij, sizeof(double) * (size_t)N * n_terms, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->KQ.p, fft_kernel_tilde, sizeof(double) * HH, hipMemcpyHostToDevice));
    TRY(ensure(c->box, sizeof(int) * N));
    TRY(ensure(c->wx, sizeof(double) * (size_t)N * p));
    TRY(ensure(c->wy, sizeof(double) * (size_t)N * p));
    launch(c, "weights_f64", pick_weights<2>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->xs.as<const double>(),
           c->ys.as<const double>(), N, g, c->mean.as<const double>(), c->mean.as<const double>() + n_boxes,
           c->box.as<int>(), c->wx.as<double>(), c->wy.as<double>());
    const long nodes = (long)g.n1d * g.n1d;
    const int nt = n_terms;
    TRY(ensure(c->w, sizeof(double) * nodes * nt));
    TRY(ensure(c->V, sizeof(double) * nodes * nt));
    launch(c, "fill", k_fill<double>, dim3(std::min(blocks_for(nodes * nt), 2048)), dim3(kThreads), 0,
           c->w.as<double>(), nodes * nt, 0.0);
    launch(c, "spread_f64", pick_spread_f64<2>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->ch.as<const double>(), nt, N, g, c->w.as<double>());
    const size_t plane = (size_t)g.H * g.n1d;
    TRY(ensure(c->spec, sizeof(cx<double>) * plane * nt));
    TRY(ensure(c->ospec, sizeof(cx<double>) * plane * nt));
    const int gr_rows = 8 * ((g.n1d + 7) / 8), gr_cols = 8 * ((g.H + 7) / 8);
    const int lines = (nt + 1) / 2;
    launch(c, "rows_fwd64", k_rows_fwd<double, kRowThreads, kRowTot>, dim3(gr_rows), dim3(kRowThreads),
           sizeof(cx<double>) * (lines + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->w.as<const double>(), nt, nt, c->spec.as<cx<double>>(),
           (unsigned long long *)nullptr);
    const int cl = std::min(nt, kColTot64 / g.Mp);
    launch(c, "cols_plain64", k_cols_plain<double, kColThreads64, kColTot64>, dim3(gr_cols), dim3(kColThreads64),
           sizeof(cx<double>) * (cl + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->spec.as<const cx<double>>(), nt, c->KQ.as<const double>(),
           c->ospec.as<cx<double>>());
    launch(c, "rows_inv64", k_rows_inv<double, kRowThreads, kRowTot>, dim3(gr_rows), dim3(kRowThreads),
           sizeof(cx<double>) * (lines + 1) * g.Mp, c->dgeo.as<const DevGeo>(), tw, c->ospec.as<const cx<double>>(), nt,
           c->V.as<double>(), nt, (const double *)nullptr, 0, 0.0, (double *)nullptr);
    launch(c, "interp_f64", pick_interp_f64<2>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->V.as<const double>(), nt, N, g, c->potA.as<double>());
    HIPTRY(hipMemcpyAsync(potentialQij, c->potA.p, sizeof(double) * (size_t)N * n_terms, hipMemcpyDeviceToHost,
                          c->stream));
    c->N = 0;
    return sync_ctx(c);
}

extern "C" int fitsne_precompute_1d(double y_min, double y_max, int n_boxes, int n_interpolation_points, int kernel,
                                    double *box_lower_bounds, double *box_upper_bounds, double *y_tilde_spacing,
                                    double *y_tilde, double *fft_kernel_vector, double df) {
    if (n_boxes < 1 || !kernel_kind_ok(kernel) || !box_lower_bounds || !box_upper_bounds || !y_tilde_spacing ||
        !y_tilde || !fft_kernel_vector)
        return fail(FITSNE_E_ARG, "invalid precompute arguments");
    const int p = n_interpolation_points;
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    const double box_width = (y_max - y_min) / (double)n_boxes;
    for (int b = 0; b < n_boxes; b++) {
        box_lower_bounds[b] = b * box_width + y_min;
        box_upper_bounds[b] = (b + 1) * box_width + y_min;
    }
    double h = 1 / (double)p;
    y_tilde_spacing[0] = h / 2;
    for (int i = 1; i < p; i++) y_tilde_spacing[i] = y_tilde_spacing[i - 1] + h;
    const int total = p * n_boxes;
    h = h * box_width;
    y_tilde[0] = y_min + h / 2;
    for (int i = 1; i < total; i++) y_tilde[i] = y_tilde[i - 1] + h;
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, p, 1.0, 1, df));
    HIPTRY(hipSetDevice(c->device));
    Geo g = mirror_geo(1, p, n_boxes, y_min, box_width, nullptr, nullptr);
    if (g.Mp > kMaxTot64) return fail(FITSNE_E_UNSUPPORTED, "grid too large for this build");
    const cx<double> *tw;
    TRY(twiddles<double>(c, g.Mp, &tw));
    TRY(put_geo(c, g, true));
    TRY(ensure(c->KQ, sizeof(double) * g.Mp));
    launch(c, "kspec1d", k_kspec1d<double, kConv1dThreads, kMaxTot64>, dim3(1), dim3(kConv1dThreads),
           sizeof(cx<double>) * g.Mp, c->dgeo.as<const DevGeo>(), tw, kernel, df, c->KQ.as<double>());
    HIPTRY(hipMemcpyAsync(fft_kernel_vector, c->KQ.p, sizeof(double) * g.Mp, hipMemcpyDeviceToHost, c->stream));
    return sync_ctx(c);
}

extern "C" int fitsne_n_body_fft_1d(int N, int n_terms, const double *Y, const double *chargesQij, int n_boxes,
                                    int n_interpolation_points, const double *box_lower_bounds,
                                    const double *box_upper_bounds, const double *y_tilde_spacings,
                                    const double *y_tilde, const double *fft_kernel_vector, double *potentialsQij) {
    (void)y_tilde;
    if (N <= 0 || n_terms < 1 || n_boxes < 1 || !Y || !chargesQij || !box_lower_bounds || !box_upper_bounds ||
        !y_tilde_spacings || !fft_kernel_vector || !potentialsQij)
        return fail(FITSNE_E_ARG, "invalid nbodyfft arguments");
    if (n_terms > 4) return fail(FITSNE_E_UNSUPPORTED, "n_terms must be in [1, 4] in this build");
    const int p = n_interpolation_points;
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, p, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    Geo g = mirror_geo(1, p, n_boxes, box_lower_bounds[0], box_upper_bounds[0] - box_lower_bounds[0],
                       box_lower_bounds, box_upper_bounds);
    for (int i = 0; i < p; i++) g.sp[i] = y_tilde_spacings[i];
    for (int i = 0; i < p; i++) {
        g.den[i] = 1;
        for (int j = 0; j < p; j++)
            if (i != j) g.den[i] *= g.sp[i] - g.sp[j];
    }
    if (g.Mp > kMaxTot64) return fail(FITSNE_E_UNSUPPORTED, "grid too large for this build");
    const cx<double> *tw;
    TRY(twiddles<double>(c, g.Mp, &tw));
    TRY(put_geo(c, g, true));
    const int nt = n_terms;
    TRY(ensure(c->xs, sizeof(double) * N));
    TRY(ensure(c->mean, sizeof(double) * n_boxes));
    TRY(ensure(c->ch, sizeof(double) * (size_t)N * nt));
    TRY(ensure(c->potA, sizeof(double) * (size_t)N * nt));
    TRY(ensure(c->KQ, sizeof(double) * g.Mp));
    HIPTRY(hipMemcpy(c->xs.p, Y, sizeof(double) * N, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->mean.p, box_lower_bounds, sizeof(double) * n_boxes, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->ch.p, chargesQij, sizeof(double) * (size_t)N * nt, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->potA.p, potentialsQij, sizeof(double) * (size_t)N * nt, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->KQ.p, fft_kernel_vector, sizeof(double) * g.Mp, hipMemcpyHostToDevice));
    TRY(ensure(c->box, sizeof(int) * N));
    TRY(ensure(c->wx, sizeof(double) * (size_t)N * p));
    TRY(ensure(c->wy, sizeof(double) * (size_t)N * p));
    launch(c, "weights_f64", pick_weights<1>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->xs.as<const double>(),
           (const double *)nullptr, N, g, c->mean.as<const double>(), (const double *)nullptr, c->box.as<int>(),
           c->wx.as<double>(), c->wy.as<double>());
    TRY(ensure(c->w, sizeof(double) * (size_t)g.n1d * nt));
    TRY(ensure(c->V, sizeof(double) * (size_t)g.n1d * nt));
    launch(c, "fill", k_fill<double>, dim3(std::min(blocks_for((long)g.n1d * nt), 2048)), dim3(kThreads), 0,
           c->w.as<double>(), (long)g.n1d * nt, 0.0);
    launch(c, "spread_f64", pick_spread_f64<1>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->ch.as<const double>(), nt, N, g, c->w.as<double>());
    launch(c, "conv1d_f64", k_conv1d<double, kConv1dThreads, kMaxTot64>, dim3(1), dim3(kConv1dThreads),
           sizeof(cx<double>) * g.Mp, c->dgeo.as<const DevGeo>(), tw, 0, 0, 1.0, 0, c->w.as<const double>(), nt, nt, 0, c->V.as<double>(),
           nt, (double *)nullptr, c->KQ.as<const double>(), (const double *)nullptr);
    launch(c, "interp_f64", pick_interp_f64<1>(p), dim3(blocks_for(N)), dim3(kThreads), 0, c->box.as<const int>(),
           c->wx.as<const double>(), c->wy.as<const double>(), c->V.as<const double>(), nt, N, g, c->potA.as<double>());
    HIPTRY(hipMemcpyAsync(potentialsQij, c->potA.p, sizeof(double) * (size_t)N * nt, hipMemcpyDeviceToHost,
                          c->stream));
    return sync_ctx(c);
}

extern "C" int fitsne_interpolate(int n_interpolation_points, int N, const double *y_in_box,
                                  const double *y_tilde_spacings, double *interpolated_values) {
    const int p = n_interpolation_points;
    if (N <= 0 || !y_in_box || !y_tilde_spacings || !interpolated_values) return fail(FITSNE_E_ARG, "invalid arguments");
    if (p < 2 || p > kMaxP) return fail(FITSNE_E_UNSUPPORTED, "n_interpolation_points must be in [2, 5]");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, p, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    Geo g{};
    for (int i = 0; i < p; i++) g.sp[i] = y_tilde_spacings[i];
    for (int i = 0; i < p; i++) {
        g.den[i] = 1;
        for (int j = 0; j < p; j++)
            if (i != j) g.den[i] *= g.sp[i] - g.sp[j];
    }
    TRY(ensure(c->xs, sizeof(double) * N));
    TRY(ensure(c->wx, sizeof(double) * (size_t)N * p));
    HIPTRY(hipMemcpy(c->xs.p, y_in_box, sizeof(double) * N, hipMemcpyHostToDevice));
    auto k = p == 2 ? k_lagrange_f64<2> : p == 3 ? k_lagrange_f64<3> : p == 4 ? k_lagrange_f64<4> : k_lagrange_f64<5>;
    launch(c, "lagrange", k, dim3(blocks_for(N)), dim3(kThreads), 0, c->xs.as<const double>(), N, g, c->wx.as<double>());
    HIPTRY(hipMemcpyAsync(interpolated_values, c->wx.p, sizeof(double) * (size_t)N * p, hipMemcpyDeviceToHost,
                          c->stream));
    return sync_ctx(c);
}

extern "C" int fitsne_evaluate_error_fft(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                                         const double *Y, int N, int D, double sum_Q, double df, double *C) {
    if (!row_P || !col_P || !val_P || !Y || !C || N <= 0 || (D != 1 && D != 2))
        return fail(FITSNE_E_ARG, "invalid evaluate_error_fft arguments");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, D, 3, 1.0, 1, df));
    TRY(set_P_impl(c, N, row_P, col_P, val_P));
    TRY(fitsne_set_Y(c, Y));
    TRY(ensure(c->zbuf, sizeof(double) * 4));
    double z[2] = {sum_Q, 1.0 / sum_Q};
    HIPTRY(hipMemcpy(c->zbuf.p, z, sizeof(z), hipMemcpyHostToDevice));
    c->have_grad = true;
    return fitsne_kl(c, C);
}

// ---------------------------------------------------------------- KL without allocation
// evaluateErrorFft for callers that cannot allocate device memory: the CSR, Y, Z and the block partials
// live in one file-scope device array (`scratch`, kNoAllocScratch bytes per device, held until the process
// ends).  Host code never passes the array's name to a kernel or a copy: it takes the device address from
// hipGetSymbolAddress once per device and uses that.  Calls are serialised by a process-wide lock; two
// processes have separate arrays.  Nothing in the array needs to start at zero: every word a call reads was
// written earlier in that call.
constexpr size_t kNoAllocScratch = 32ull << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocMaxDevices];

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" int fitsne_evaluate_error_fft_noalloc(const uint32_t *row_P, const uint32_t *col_P, const double *val_P,
                                                 const double *Y, int N, int D, double sum_Q, double df, double *C) {
    if (!row_P || !col_P || !val_P || !Y || !C || N <= 0 || (D != 1 && D != 2))
        return fail(FITSNE_E_ARG, "invalid evaluate_error_fft arguments");
    const long long nnz = (long long)row_P[N];
    if (row_P[0] != 0) return fail(FITSNE_E_ARG, "row_P must start at 0");
    for (int i = 0; i < N; i++)
        if (row_P[i + 1] < row_P[i]) return fail(FITSNE_E_ARG, "row_P must be non-decreasing");
    for (long long k = 0; k < nnz; k++)
        if (col_P[k] >= (uint32_t)N) return fail(FITSNE_E_ARG, "col_P entry out of range");
    for (size_t e = 0; e < (size_t)N * D; e++)
        if (!std::isfinite(Y[e])) return fail(FITSNE_E_ARG, "Y contains non-finite values");
    // the same launch as the context path (32 lanes per row), so the partials and their sum are the same
    const long threads = (long)N * 32;
    const int nblk = blocks_for(threads);
    const size_t o_row = 0, o_col = o_row + align256(sizeof(uint32_t) * ((size_t)N + 1));
    const size_t o_val = o_col + align256(sizeof(uint32_t) * (size_t)std::max<long long>(nnz, 1));
    const size_t o_y = o_val + align256(sizeof(double) * (size_t)std::max<long long>(nnz, 1));
    const size_t o_z = o_y + align256(sizeof(double) * (size_t)N * D);
    const size_t o_part = o_z + align256(sizeof(double) * 2);
    const size_t o_out = o_part + align256(sizeof(double) * (size_t)nblk);
    const size_t need = o_out + align256(sizeof(double) * 2);
    if (need > kNoAllocScratch)
        return fail(FITSNE_E_ARG, "evaluate_error_fft_noalloc: N = " + std::to_string(N) + ", nnz = " +
                                      std::to_string(nnz) + " need " + std::to_string(need) +
                                      " bytes of device scratch; this entry point holds " +
                                      std::to_string(kNoAllocScratch) + " (use fitsne_evaluate_error_fft)");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    int dev = 0;
    HIPTRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(FITSNE_E_UNSUPPORTED, "device index beyond this build's table");
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        HIPTRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(FITSNE_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_noalloc_ptr[dev] = ptr;
    }
    unsigned char *base = static_cast<unsigned char *>(g_noalloc_ptr[dev]);
    auto *rp = reinterpret_cast<uint32_t *>(base + o_row);
    auto *cp = reinterpret_cast<uint32_t *>(base + o_col);
    auto *vp = reinterpret_cast<double *>(base + o_val);
    auto *yp = reinterpret_cast<double *>(base + o_y);
    auto *zp = reinterpret_cast<double *>(base + o_z);
    auto *part = reinterpret_cast<double *>(base + o_part);
    auto *out = reinterpret_cast<double *>(base + o_out);
    const double z[2] = {sum_Q, 1.0 / sum_Q};
    HIPTRY(hipMemcpy(rp, row_P, sizeof(uint32_t) * ((size_t)N + 1), hipMemcpyHostToDevice));
    if (nnz) {
        HIPTRY(hipMemcpy(cp, col_P, sizeof(uint32_t) * (size_t)nnz, hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(vp, val_P, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
    }
    HIPTRY(hipMemcpy(yp, Y, sizeof(double) * (size_t)N * D, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(zp, z, sizeof(z), hipMemcpyHostToDevice));
    if (D == 2)
        k_kl<double, uint32_t, uint32_t, 2, 32><<<dim3(nblk), dim3(kThreads), 0, 0>>>(rp, cp, vp, 1.0, yp, N, df, zp, part);
    else
        k_kl<double, uint32_t, uint32_t, 1, 32><<<dim3(nblk), dim3(kThreads), 0, 0>>>(rp, cp, vp, 1.0, yp, N, df, zp, part);
    k_sum_final<<<dim3(1), dim3(kThreads), 0, 0>>>(part, nblk, 0.0, out);
    HIPTRY(hipGetLastError());
    double res[2];
    HIPTRY(hipMemcpy(res, out, sizeof(res), hipMemcpyDeviceToHost));
    *C = res[0];
    return FITSNE_OK;
}

extern "C" int fitsne_distances2similarities_batch(const double *D, double *P, int n_points, int K,
                                                   double perplexity, double sigma, int ifSquared,
                                                   const double *beta0, double *beta, int *iters) {
    if (!D || !P || !beta || n_points <= 0 || K <= 0) return fail(FITSNE_E_ARG, "invalid arguments");
    if (!(perplexity > 0) && !(sigma > 0)) return fail(FITSNE_E_ARG, "need perplexity > 0 or sigma > 0");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, 3, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    const size_t nk = (size_t)n_points * K;
    TRY(ensure(c->w, sizeof(double) * nk));
    TRY(ensure(c->V, sizeof(double) * nk));
    TRY(ensure(c->xs, sizeof(double) * n_points));
    TRY(ensure(c->ys, sizeof(double) * n_points));
    TRY(ensure(c->box, sizeof(int) * n_points));
    HIPTRY(hipMemcpy(c->w.p, D, sizeof(double) * nk, hipMemcpyHostToDevice));
    if (beta0) HIPTRY(hipMemcpy(c->ys.p, beta0, sizeof(double) * n_points, hipMemcpyHostToDevice));
    const long threads = (long)n_points * kWave;
    launch(c, "d2s", k_d2s, dim3(blocks_for(threads)), dim3(kThreads), 0, c->w.as<const double>(), c->V.as<double>(),
           n_points, K, perplexity, sigma, ifSquared, beta0 ? c->ys.as<const double>() : (const double *)nullptr,
           c->xs.as<double>(), c->box.as<int>(), 0, 0);
    HIPTRY(hipMemcpyAsync(P, c->V.p, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(beta, c->xs.p, sizeof(double) * n_points, hipMemcpyDeviceToHost, c->stream));
    if (iters) HIPTRY(hipMemcpyAsync(iters, c->box.p, sizeof(int) * n_points, hipMemcpyDeviceToHost, c->stream));
    return sync_ctx(c);
}

extern "C" int fitsne_symmetrize(const uint32_t *row_P, const uint32_t *col_P, const double *val_P, int N,
                                 uint32_t *sym_row_P, uint32_t *sym_col_P, double *sym_val_P, long long *sym_nnz) {
    if (!row_P || !col_P || !val_P || !sym_row_P || !sym_col_P || !sym_val_P || N <= 0)
        return fail(FITSNE_E_ARG, "invalid symmetrisation arguments");
    if (row_P[0] != 0) return fail(FITSNE_E_ARG, "row_P must start at 0");
    for (int i = 0; i < N; i++)
        if (row_P[i + 1] < row_P[i]) return fail(FITSNE_E_ARG, "row_P must be non-decreasing");
    const long nnz = (long)row_P[N];
    if (nnz > (long)(INT32_MAX / 2)) return fail(FITSNE_E_UNSUPPORTED, "more than 2^30 entries");
    for (long k = 0; k < nnz; k++)
        if (col_P[k] >= (uint32_t)N) return fail(FITSNE_E_ARG, "col_P entry out of range");
    const int cbits = bits_for((unsigned long long)N - 1), rbits = bits_for((unsigned long long)N);
    const int ebits = bits_for((unsigned long long)std::max<long>(nnz - 1, 0));
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, 3, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    if (nnz == 0) {
        for (int i = 0; i <= N; i++) sym_row_P[i] = 0;
        if (sym_nnz) *sym_nnz = 0;
        return FITSNE_OK;
    }
    const size_t ne = (size_t)nnz;
    TRY(ensure(c->sy_row, sizeof(uint32_t) * (N + 1)));
    TRY(ensure(c->sy_col, sizeof(uint32_t) * ne));
    TRY(ensure(c->sy_val, sizeof(double) * ne));
    TRY(ensure(c->sy_erow, sizeof(int) * ne));
    TRY(ensure(c->sy_keys, sizeof(uint64_t) * ne));
    TRY(ensure(c->sy_keys2, sizeof(uint64_t) * ne));
    TRY(ensure(c->sy_vals, sizeof(uint32_t) * ne));
    TRY(ensure(c->sy_vals2, sizeof(uint32_t) * ne));
    TRY(ensure(c->sy_partner, sizeof(int64_t) * ne));
    TRY(ensure(c->sy_cnt, sizeof(int) * N));
    TRY(ensure(c->sy_start, sizeof(int) * (N + 1)));
    TRY(ensure(c->sy_okeys, sizeof(uint64_t) * 2 * ne));
    TRY(ensure(c->sy_okeys2, sizeof(uint64_t) * 2 * ne));
    TRY(ensure(c->sy_scol, sizeof(uint32_t) * 2 * ne));
    TRY(ensure(c->sy_sval, sizeof(double) * 2 * ne));
    HIPTRY(hipMemcpyAsync(c->sy_row.p, row_P, sizeof(uint32_t) * (N + 1), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemcpyAsync(c->sy_col.p, col_P, sizeof(uint32_t) * ne, hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemcpyAsync(c->sy_val.p, val_P, sizeof(double) * ne, hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemsetAsync(c->sy_cnt.p, 0, sizeof(int) * N, c->stream));
    const int gblk = std::min(blocks_for(nnz), 8192);
    launch(c, "sym_rows", k_entry_rows<uint32_t>, dim3(blocks_for((long)N * kWave)), dim3(kThreads), 0,
           c->sy_row.as<const uint32_t>(), N, c->sy_erow.as<int>());
    launch(c, "sym_keys", k_sym_keys, dim3(gblk), dim3(kThreads), 0, c->sy_erow.as<const int>(),
           c->sy_col.as<const uint32_t>(), nnz, cbits, c->sy_keys.as<uint64_t>(), c->sy_vals.as<uint32_t>());
    // radix sorts: temporary storage sized for the larger of the two
    size_t t1 = 0, t2 = 0;
    HIPTRY(rocprim::radix_sort_pairs(nullptr, t1, c->sy_keys.as<uint64_t>(), c->sy_keys2.as<uint64_t>(),
                                     c->sy_vals.as<uint32_t>(), c->sy_vals2.as<uint32_t>(), ne, 0u,
                                     (unsigned)(rbits + cbits), c->stream));
    HIPTRY(rocprim::radix_sort_keys(nullptr, t2, c->sy_okeys.as<uint64_t>(), c->sy_okeys2.as<uint64_t>(), 2 * ne, 0u,
                                    (unsigned)(rbits + ebits), c->stream));
    TRY(ensure(c->sy_tmp, std::max<size_t>(std::max(t1, t2), 16)));
    HIPTRY(rocprim::radix_sort_pairs(c->sy_tmp.p, t1, c->sy_keys.as<uint64_t>(), c->sy_keys2.as<uint64_t>(),
                                     c->sy_vals.as<uint32_t>(), c->sy_vals2.as<uint32_t>(), ne, 0u,
                                     (unsigned)(rbits + cbits), c->stream));
    const uint64_t none = (uint64_t)N << ebits;
    launch(c, "sym_place", k_sym_place<uint32_t>, dim3(gblk), dim3(kThreads), 0, c->sy_row.as<const uint32_t>(),
           c->sy_erow.as<const int>(), c->sy_col.as<const uint32_t>(), c->sy_keys2.as<const uint64_t>(),
           c->sy_vals2.as<const uint32_t>(), nnz, cbits, ebits, none, c->sy_partner.as<int64_t>(),
           c->sy_cnt.as<int>(), c->sy_okeys.as<uint64_t>());
    HIPTRY(rocprim::radix_sort_keys(c->sy_tmp.p, t2, c->sy_okeys.as<uint64_t>(), c->sy_okeys2.as<uint64_t>(), 2 * ne,
                                    0u, (unsigned)(rbits + ebits), c->stream));
    launch(c, "scan_counts", k_scan_counts, dim3(1), dim3(1024), 0, c->sy_cnt.as<int>(), c->sy_start.as<int>(), N);
    launch(c, "sym_fill", k_sym_fill, dim3(std::min(blocks_for(2 * nnz), 8192)), dim3(kThreads), 0,
           c->sy_okeys2.as<const uint64_t>(), c->sy_start.as<const int>() + N, ebits, c->sy_erow.as<const int>(),
           c->sy_col.as<const uint32_t>(), c->sy_val.as<const double>(), c->sy_partner.as<const int64_t>(),
           c->sy_scol.as<uint32_t>(), c->sy_sval.as<double>());
    HIPTRY(hipGetLastError());
    std::vector<int> start((size_t)N + 1);
    HIPTRY(hipMemcpyAsync(start.data(), c->sy_start.p, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
    TRY(sync_ctx(c));
    const long M = start[N];
    for (int i = 0; i <= N; i++) sym_row_P[i] = (uint32_t)start[i];
    HIPTRY(hipMemcpy(sym_col_P, c->sy_scol.p, sizeof(uint32_t) * (size_t)M, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(sym_val_P, c->sy_sval.p, sizeof(double) * (size_t)M, hipMemcpyDeviceToHost));
    if (sym_nnz) *sym_nnz = M;
    return FITSNE_OK;
}

extern "C" int fitsne_knn_exact(const double *X, int N, int D, int K, int32_t *idx, double *dist) {
    if (!X || !idx || !dist || N < 2 || D < 1 || K < 1) return fail(FITSNE_E_ARG, "invalid kNN arguments");
    if (K > N - 1) return fail(FITSNE_E_ARG, "K must be at most N - 1");
    if (K > 1000) return fail(FITSNE_E_UNSUPPORTED, "K above 1000 is not supported");
    const size_t nd = (size_t)N * D;
    for (size_t i = 0; i < nd; i++)
        if (!std::isfinite(X[i])) return fail(FITSNE_E_ARG, "X contains non-finite values");
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, 3, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    TRY(ensure(c->knnX, sizeof(double) * nd));
    TRY(ensure(c->knnXT, sizeof(double) * nd));
    TRY(ensure(c->knnI, sizeof(int) * (size_t)N * K));
    TRY(ensure(c->knnD, sizeof(double) * (size_t)N * K));
    HIPTRY(hipMemcpy(c->knnX.p, X, sizeof(double) * nd, hipMemcpyHostToDevice));
    launch(c, "transpose64", k_transpose_f64, dim3((N + 31) / 32, (D + 31) / 32), dim3(256), 0,
           c->knnX.as<const double>(), (long)N, D, c->knnXT.as<double>());
    int B = 128;
    while (B < 2 * K || B < K + kWave) B *= 2;
    const size_t lds = (size_t)kKnnQ * B * (sizeof(double) + sizeof(int));
    launch(c, "knn64", k_knn_f64, dim3((N + kKnnQ - 1) / kKnnQ), dim3(kWave), lds, c->knnX.as<const double>(),
           c->knnXT.as<const double>(), N, D, K, B, c->knnI.as<int>(), c->knnD.as<double>(), (const int *)nullptr, N);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(idx, c->knnI.p, sizeof(int) * (size_t)N * K, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(dist, c->knnD.p, sizeof(double) * (size_t)N * K, hipMemcpyDeviceToHost, c->stream));
    return sync_ctx(c);
}

// ---------------------------------------------------------------- kNN for large N (FP32 MFMA candidates)

// Device buffers of one call, released on every return.
struct KnnmBufs {
    DBuf X, XT, XT4, nrm, nrm64, XS4, nrs, tau, cand, cnt, idx, dist, redo, nredo, nmax;
    ~KnnmBufs() {
        DBuf *all[] = {&X, &XT, &XT4, &nrm, &nrm64, &XS4, &nrs, &tau, &cand, &cnt, &idx, &dist, &redo, &nredo, &nmax};
        for (DBuf *b : all) release(*b);
    }
};

template <int NCH>
static void knnm_passes(hipStream_t st, KnnmBufs &m, int N, int NS) {
    const int nblk = (N + kKmWaves * kKmQ - 1) / (kKmWaves * kKmQ);
    k_knnm_tau<NCH><<<dim3(nblk), dim3(kKmWaves * kWave), 0, st>>>(m.XT4.as<const float>(), N, m.nrm.as<const float>(),
                                                                   m.XS4.as<const float>(), NS,
                                                                   m.nrs.as<const float>(), m.tau.as<float>());
    k_knnm_collect<NCH><<<dim3(nblk), dim3(kKmWaves * kWave), 0, st>>>(m.XT4.as<const float>(), N,
                                                                       m.nrm.as<const float>(),
                                                                       m.tau.as<const float>(), m.cand.as<int>(),
                                                                       m.cnt.as<int>());
}

extern "C" int fitsne_knn_mfma(const double *X, int N, int D, int K, int32_t *idx, double *dist,
                               long long *recomputed) {
    if (!X || !idx || !dist || N < 2 || D < 1 || K < 1) return fail(FITSNE_E_ARG, "invalid kNN arguments");
    if (K > N - 1) return fail(FITSNE_E_ARG, "K must be at most N - 1");
    if (K > 1000) return fail(FITSNE_E_UNSUPPORTED, "K above 1000 is not supported");
    if (D > 128) return fail(FITSNE_E_UNSUPPORTED, "the MFMA kNN takes D <= 128 (use fitsne_knn_exact)");
    if (K > kKmCap / 2) return fail(FITSNE_E_UNSUPPORTED, "the MFMA kNN takes K <= 256 (use fitsne_knn_exact)");
    const size_t nd = (size_t)N * D;
    for (size_t i = 0; i < nd; i++)
        if (!std::isfinite(X[i])) return fail(FITSNE_E_ARG, "X contains non-finite values");
    const int need = (D + 3) / 4;
    const int nch = need <= 8 ? 8 : need <= 13 ? 13 : need <= 16 ? 16 : need <= 25 ? 25 : 32;
    const int NS = (N + kKmStride - 1) / kKmStride;
    std::lock_guard<std::mutex> lk(g_mirror_mu);
    fitsne_ctx *c;
    TRY(mirror_ctx(&c, 1, 3, 1.0, 1, 1.0));
    HIPTRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    KnnmBufs m;
    TRY(ensure(m.X, sizeof(double) * nd));
    TRY(ensure(m.XT4, sizeof(float4) * (size_t)nch * N));
    TRY(ensure(m.nrm, sizeof(float) * N));
    TRY(ensure(m.nrm64, sizeof(double) * N));
    TRY(ensure(m.XS4, sizeof(float4) * (size_t)nch * NS));
    TRY(ensure(m.nrs, sizeof(float) * NS));
    TRY(ensure(m.tau, sizeof(float) * N));
    TRY(ensure(m.cand, sizeof(int) * (size_t)N * kKmCap));
    TRY(ensure(m.cnt, sizeof(int) * N));
    TRY(ensure(m.idx, sizeof(int) * (size_t)N * K));
    TRY(ensure(m.dist, sizeof(double) * (size_t)N * K));
    TRY(ensure(m.redo, sizeof(int) * N));
    TRY(ensure(m.nredo, sizeof(int)));
    TRY(ensure(m.nmax, sizeof(double)));
    HIPTRY(hipMemcpy(m.X.p, X, sizeof(double) * nd, hipMemcpyHostToDevice));
    k_fill<int><<<dim3(1), dim3(kWave), 0, st>>>(m.nredo.as<int>(), 1, 0);
    k_knnm_prep<<<dim3((N + 255) / 256), dim3(256), 0, st>>>(m.X.as<const double>(), N, D, nch, m.XT4.as<float4>(),
                                                             m.nrm.as<float>(), m.nrm64.as<double>(), NS,
                                                             m.XS4.as<float4>(), m.nrs.as<float>());
    k_knnm_max<<<dim3(1), dim3(1024), 0, st>>>(m.nrm64.as<const double>(), (long)N, m.nmax.as<double>());
    switch (nch) {
    case 8: knnm_passes<8>(st, m, N, NS); break;
    case 13: knnm_passes<13>(st, m, N, NS); break;
    case 16: knnm_passes<16>(st, m, N, NS); break;
    case 25: knnm_passes<25>(st, m, N, NS); break;
    default: knnm_passes<32>(st, m, N, NS); break;
    }
    const double epsu = 2.0 * (4.0 * nch + 8.0) * std::ldexp(1.0, -24);
    k_knnm_exact<<<dim3(N), dim3(kWave), 0, st>>>(m.X.as<const double>(), N, D, K, m.cand.as<const int>(),
                                                  m.cnt.as<const int>(), m.tau.as<const float>(),
                                                  m.nrm64.as<const double>(), m.nmax.as<const double>(), epsu,
                                                  m.idx.as<int>(), m.dist.as<double>(), m.redo.as<int>(),
                                                  m.nredo.as<int>());
    HIPTRY(hipGetLastError());
    int nredo = 0;
    HIPTRY(hipMemcpyAsync(&nredo, m.nredo.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    if (nredo < 0 || nredo > N) return fail(FITSNE_E_HIP, "inconsistent recompute count");
    if (nredo > 0) {  // the queries the candidate lists could not settle: the FP64 kernel over all points
        TRY(ensure(m.XT, sizeof(double) * nd));
        k_transpose_f64<<<dim3((N + 31) / 32, (D + 31) / 32), dim3(256), 0, st>>>(m.X.as<const double>(), (long)N, D,
                                                                                m.XT.as<double>());
        int B = 128;
        while (B < 2 * K || B < K + kWave) B *= 2;
        const size_t lds = (size_t)kKnnQ * B * (sizeof(double) + sizeof(int));
        k_knn_f64<<<dim3((nredo + kKnnQ - 1) / kKnnQ), dim3(kWave), lds, st>>>(
            m.X.as<const double>(), m.XT.as<const double>(), N, D, K, B, m.idx.as<int>(), m.dist.as<double>(),
            m.redo.as<const int>(), nredo);
        HIPTRY(hipGetLastError());
    }
    HIPTRY(hipMemcpyAsync(idx, m.idx.p, sizeof(int) * (size_t)N * K, hipMemcpyDeviceToHost, st));
    HIPTRY(hipMemcpyAsync(dist, m.dist.p, sizeof(double) * (size_t)N * K, hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    if (recomputed) *recomputed = nredo;
    return FITSNE_OK;
}

// ============================================================================ learned β warm start

struct fitsne_beta_model {
    int device = 0, K = 0;
    hipStream_t stream = nullptr;
    DBuf prm, m1, m2, grad, gpart, lpart, w1t, w2t, w2, X, tgt, dist, beta0;
    long steps = 0;
};

namespace {

uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double gauss(uint64_t &st) {
    const double u1 = ((splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int mlp_sync(fitsne_beta_model *m) {
    HIPTRY(hipStreamSynchronize(m->stream));
    return FITSNE_OK;
}

// features of dist[n][K] into m->X (bf16 [n][96]) with the model's mu / sd
int mlp_features(fitsne_beta_model *m, const double *dist, long n) {
    TRY(ensure(m->dist, sizeof(double) * (size_t)n * m->K));
    TRY(ensure(m->X, sizeof(unsigned short) * (size_t)n * kMlpIn));
    HIPTRY(hipMemcpyAsync(m->dist.p, dist, sizeof(double) * (size_t)n * m->K, hipMemcpyHostToDevice, m->stream));
    const long tot = n * kMlpIn;
    k_mlp_features<<<dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads), 0, m->stream>>>(
        m->dist.as<const double>(), (int)n, m->K, m->prm.as<const float>(), m->X.as<unsigned short>());
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

int mlp_pack(fitsne_beta_model *m) {
    k_mlp_pack<<<dim3((kMlpHid * kMlpIn + kThreads - 1) / kThreads), dim3(kThreads), 0, m->stream>>>(
        m->prm.as<const float>(), m->w1t.as<unsigned short>(), m->w2t.as<unsigned short>(), m->w2.as<unsigned short>());
    HIPTRY(hipGetLastError());
    return FITSNE_OK;
}

// gradient over X rows [off, off + n) with targets tgt[off ..]; optional Adam step
int mlp_grad_step(fitsne_beta_model *m, long off, long n, bool step, float lr, int *nparts_out) {
    const long ntiles = (n + 63) / 64;
    const int G = (int)std::min<long>(ntiles, 256);
    TRY(ensure(m->gpart, sizeof(float) * (size_t)G * kMlpTrainN));
    TRY(ensure(m->lpart, sizeof(double) * (size_t)G));
    TRY(mlp_pack(m));
    k_mlp_grad<<<dim3(G), dim3(kMlpThreads), 0, m->stream>>>(
        m->X.as<const unsigned short>() + off * kMlpIn, m->tgt.as<const float>() + off, n, 1.0f / (float)n,
        m->w1t.as<const unsigned short>(), m->w2t.as<const unsigned short>(), m->w2.as<const unsigned short>(),
        m->prm.as<const float>(), m->gpart.as<float>(), m->lpart.as<double>());
    HIPTRY(hipGetLastError());
    float c1 = 1.f, c2 = 1.f;
    if (step) {
        m->steps++;
        c1 = (float)(1.0 - std::pow(0.9, (double)m->steps));
        c2 = (float)(1.0 - std::pow(0.999, (double)m->steps));
    }
    k_mlp_reduce_adam<<<dim3((kMlpTrainN + kThreads - 1) / kThreads), dim3(kThreads), 0, m->stream>>>(
        m->gpart.as<const float>(), G, m->grad.as<float>(), m->prm.as<float>(), m->m1.as<float>(), m->m2.as<float>(),
        step ? 1 : 0, lr, c1, c2);
    HIPTRY(hipGetLastError());
    if (nparts_out) *nparts_out = G;
    return FITSNE_OK;
}

int mlp_loss(fitsne_beta_model *m, int G, long n, double *loss) {
    std::vector<double> lp(G);
    HIPTRY(hipMemcpyAsync(lp.data(), m->lpart.p, sizeof(double) * G, hipMemcpyDeviceToHost, m->stream));
    TRY(mlp_sync(m));
    double s = 0.0;
    for (double v : lp) s += v;
    *loss = s / (double)n;
    return FITSNE_OK;
}

}  // namespace

extern "C" int fitsne_beta_model_create(fitsne_beta_model **out, int K, unsigned long long seed) {
    if (!out) return fail(FITSNE_E_ARG, "null model pointer");
    *out = nullptr;
    if (K < 1 || K > kMlpIn) return fail(FITSNE_E_UNSUPPORTED, "the β model takes 1..96 neighbour distances");
    int n = 0;
    HIPTRY(hipGetDeviceCount(&n));
    if (n < 1) return fail(FITSNE_E_NODEVICE, "no GPU device");
    auto *m = new fitsne_beta_model();
    m->K = K;
    HIPTRY(hipSetDevice(m->device));
    HIPTRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    std::vector<float> p(kMlpParams, 0.f);
    for (int k = 0; k < kMlpIn; k++) p[kMlpOffSd + k] = 1.f;
    uint64_t st = seed;
    const double s1 = std::sqrt(2.0 / K), s2 = std::sqrt(2.0 / kMlpHid), s3 = std::sqrt(1.0 / kMlpHid);
    for (int k = 0; k < kMlpIn; k++)
        for (int j = 0; j < kMlpHid; j++) p[kMlpOffW1 + k * kMlpHid + j] = k < K ? (float)(s1 * gauss(st)) : 0.f;
    for (int e = 0; e < kMlpHid * kMlpHid; e++) p[kMlpOffW2 + e] = (float)(s2 * gauss(st));
    for (int j = 0; j < kMlpHid; j++) p[kMlpOffW3 + j] = (float)(s3 * gauss(st));
    int rc = ensure(m->prm, sizeof(float) * kMlpParams);
    if (rc == FITSNE_OK) rc = ensure(m->m1, sizeof(float) * kMlpTrainN);
    if (rc == FITSNE_OK) rc = ensure(m->m2, sizeof(float) * kMlpTrainN);
    if (rc == FITSNE_OK) rc = ensure(m->grad, sizeof(float) * kMlpTrainN);
    if (rc == FITSNE_OK) rc = ensure(m->w1t, sizeof(unsigned short) * kMlpHid * kMlpIn);
    if (rc == FITSNE_OK) rc = ensure(m->w2t, sizeof(unsigned short) * kMlpHid * kMlpHid);
    if (rc == FITSNE_OK) rc = ensure(m->w2, sizeof(unsigned short) * kMlpHid * kMlpHid);
    if (rc != FITSNE_OK) {
        fitsne_beta_model_destroy(m);
        return rc;
    }
    std::vector<float> zeros(kMlpTrainN, 0.f);
    if (hipMemcpy(m->prm.p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m1.p, zeros.data(), sizeof(float) * kMlpTrainN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m2.p, zeros.data(), sizeof(float) * kMlpTrainN, hipMemcpyHostToDevice) != hipSuccess) {
        fitsne_beta_model_destroy(m);
        return fail(FITSNE_E_HIP, "upload of the model parameters failed");
    }
    *out = m;
    return FITSNE_OK;
}

extern "C" int fitsne_beta_model_destroy(fitsne_beta_model *m) {
    if (!m) return FITSNE_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    DBuf *all[] = {&m->prm, &m->m1, &m->m2, &m->grad, &m->gpart, &m->lpart, &m->w1t, &m->w2t, &m->w2,
                   &m->X, &m->tgt, &m->dist, &m->beta0};
    for (DBuf *b : all) release(*b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return FITSNE_OK;
}

extern "C" int fitsne_beta_model_params(fitsne_beta_model *m, float *out) {
    if (!m || !out) return fail(FITSNE_E_ARG, "null argument");
    HIPTRY(hipSetDevice(m->device));
    HIPTRY(hipMemcpyAsync(out, m->prm.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    return mlp_sync(m);
}

extern "C" int fitsne_beta_model_set_params(fitsne_beta_model *m, const float *in) {
    if (!m || !in) return fail(FITSNE_E_ARG, "null argument");
    for (int k = 0; k < kMlpIn; k++)
        if (!(in[kMlpOffSd + k] > 0.f)) return fail(FITSNE_E_ARG, "feature scales must be positive");
    HIPTRY(hipSetDevice(m->device));
    HIPTRY(hipMemcpyAsync(m->prm.p, in, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    return mlp_sync(m);
}

extern "C" int fitsne_beta_model_gradient(fitsne_beta_model *m, const double *dist, const double *beta, int n,
                                          float *grad, double *loss) {
    if (!m || !dist || !beta || n <= 0) return fail(FITSNE_E_ARG, "invalid arguments");
    HIPTRY(hipSetDevice(m->device));
    std::vector<float> t(n);
    for (int s = 0; s < n; s++) {
        if (!(beta[s] > 0.0)) return fail(FITSNE_E_ARG, "targets must be positive");
        t[s] = (float)std::log(beta[s]);
    }
    TRY(mlp_features(m, dist, n));
    TRY(ensure(m->tgt, sizeof(float) * (size_t)n));
    HIPTRY(hipMemcpyAsync(m->tgt.p, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    int G = 0;
    TRY(mlp_grad_step(m, 0, n, false, 0.f, &G));
    if (grad) HIPTRY(hipMemcpyAsync(grad, m->grad.p, sizeof(float) * kMlpTrainN, hipMemcpyDeviceToHost, m->stream));
    double l = 0.0;
    TRY(mlp_loss(m, G, n, &l));
    if (loss) *loss = l;
    return FITSNE_OK;
}

extern "C" int fitsne_beta_model_fit(fitsne_beta_model *m, const double *dist, const double *beta, int n, int epochs,
                                     int batch, double lr, unsigned long long seed, double *final_loss) {
    if (!m || !dist || !beta || n < 64 || epochs < 1 || batch < 64) return fail(FITSNE_E_ARG, "invalid arguments");
    HIPTRY(hipSetDevice(m->device));
    const int K = m->K;
    // feature standardisation from the training set
    std::vector<double> mu(K, 0.0), var(K, 0.0);
    for (int s = 0; s < n; s++)
        for (int k = 0; k < K; k++) mu[k] += std::log(std::fmax(dist[(size_t)s * K + k] * dist[(size_t)s * K + k], 1e-30));
    for (int k = 0; k < K; k++) mu[k] /= n;
    for (int s = 0; s < n; s++)
        for (int k = 0; k < K; k++) {
            const double f = std::log(std::fmax(dist[(size_t)s * K + k] * dist[(size_t)s * K + k], 1e-30)) - mu[k];
            var[k] += f * f;
        }
    std::vector<float> p(kMlpParams);
    HIPTRY(hipMemcpy(p.data(), m->prm.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    double tmean = 0.0;
    for (int s = 0; s < n; s++) {
        if (!(beta[s] > 0.0)) return fail(FITSNE_E_ARG, "targets must be positive");
        tmean += std::log(beta[s]);
    }
    for (int k = 0; k < kMlpIn; k++) {
        p[kMlpOffMu + k] = k < K ? (float)mu[k] : 0.f;
        p[kMlpOffSd + k] = k < K ? (float)std::fmax(std::sqrt(var[k] / n), 1e-6) : 1.f;
    }
    p[kMlpOffB3] = (float)(tmean / n);
    HIPTRY(hipMemcpy(m->prm.p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    // one shuffled copy of the samples, then contiguous mini-batches
    std::vector<int> order(n);
    for (int s = 0; s < n; s++) order[s] = s;
    uint64_t st = seed;
    for (int s = n - 1; s > 0; s--) std::swap(order[s], order[(int)(splitmix64(st) % (uint64_t)(s + 1))]);
    std::vector<double> dsh((size_t)n * K);
    std::vector<float> t(n);
    for (int s = 0; s < n; s++) {
        std::memcpy(&dsh[(size_t)s * K], &dist[(size_t)order[s] * K], sizeof(double) * K);
        t[s] = (float)std::log(beta[order[s]]);
    }
    TRY(mlp_features(m, dsh.data(), n));
    TRY(ensure(m->tgt, sizeof(float) * (size_t)n));
    HIPTRY(hipMemcpyAsync(m->tgt.p, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    for (int e = 0; e < epochs; e++)
        for (long off = 0; off + 64 <= n; off += batch) {
            const long nb = std::min<long>(batch, n - off);
            TRY(mlp_grad_step(m, off, nb, true, (float)lr, nullptr));
        }
    int G = 0;
    TRY(mlp_grad_step(m, 0, n, false, 0.f, &G));
    double l = 0.0;
    TRY(mlp_loss(m, G, n, &l));
    if (final_loss) *final_loss = l;
    return FITSNE_OK;
}

extern "C" int fitsne_beta_model_predict(fitsne_beta_model *m, const double *dist, int n, double *beta0) {
    if (!m || !dist || !beta0 || n <= 0) return fail(FITSNE_E_ARG, "invalid arguments");
    HIPTRY(hipSetDevice(m->device));
    TRY(mlp_features(m, dist, n));
    TRY(ensure(m->beta0, sizeof(double) * (size_t)n));
    TRY(mlp_pack(m));
    const long blocks = ((long)n + 16 * kMlpWaves - 1) / (16 * kMlpWaves);
    k_mlp_predict<<<dim3((unsigned)blocks), dim3(kMlpThreads), 0, m->stream>>>(
        m->X.as<const unsigned short>(), n, m->w1t.as<const unsigned short>(), m->w2t.as<const unsigned short>(),
        m->prm.as<const float>(), m->beta0.as<double>());
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(beta0, m->beta0.p, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    return mlp_sync(m);
}
