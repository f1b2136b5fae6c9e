// fitsne_batch.hpp -- one step of many independent 2D embeddings (FP32 form) with one launch per stage.
//
// Each stage kernel covers every problem of the batch.  Stages whose work is fixed by the problem's N
// (attractive term, spread, interpolation, centring) map workgroups to problems through a prefix of
// per-problem workgroup counts computed once on the host.  The FFT passes, whose line counts follow each
// problem's grid and so are known only after the step's geometry on the device, are persistent: a
// fixed number of workgroups takes (problem, line) items from a per-stage counter, over a prefix of
// line counts that the scheduling kernel writes after the geometry.  Every line, row block and point
// group is computed by the same device code as a single context's step, so each embedding of the batch
// is bitwise the one its own context would reach.
#pragma once

#include "fitsne_kernels.hpp"

namespace fitsne {

// One problem of a batch: its records and device buffers (all owned by its context).
struct BatchProb {
    DevGeo *dg;
    const StepDev *sd;
    float *Y, *pos, *dC, *uY, *gains;
    int N;
    int spread_pts;  // points per spread workgroup
    int cblk;        // workgroups of the centring (and so min / max partials)
    long first;      // flat index of the caller's first coordinate in the internal order
    const int *inv;  // caller -> internal order (null when unpermuted), for the 2D min rule replay
    const int *wstart, *wpure;
    const uint32_t *wkey;
    const float *wval, *wbmax;
    int win_R, win_NW;
    unsigned long long *fixgrid;
    cx<float> *tw, *Trow, *spec, *ospec;
    float *KQ, *K1, *KS, *KT;
    float4 *gout;
    double *zpart, *zbuf, *part, *mean, *mmpart;
};

// FFT passes fed from per-stage work counters.  Each pass runs as two launches: one sized (LDS, and so
// occupancy) for FFT lengths up to kBatchSmallMp, which most problems of a batch have, and one for the
// rest, up to the build's limit; a problem's lines go to the class of its FFT length.  (A third class
// for lengths up to 512, with half the threads and LDS, was no faster: per complex element the passes
// run at the single-context passes' rate, and the extra launches add their latency.)
enum { kBsKernRows = 0, kBsKernCols, kBsRowsFwd, kBsCols, kBsRowsInv, kBsStages };
constexpr int kBsClasses = 2;
constexpr int kBatchSmallMp = 1024;
__host__ __device__ constexpr int batch_class(int Mp) { return Mp <= kBatchSmallMp ? 0 : 1; }
// Lines (rows or columns) per work item: as many as the class's buffers hold at FFT length Mp, so that a
// workgroup's threads and barriers serve several short lines at once and the transposed reads and writes
// of the row passes touch runs of consecutive rows.
// Short-line class buffers: lines of up to kBatchItemMp elements in all per item for the kernel-spectrum
// and row passes, kBatchSmallMp for the column pass (whose four lines per column already fill it; at
// kBatchItemMp its 82 KB of LDS left one workgroup per CU and the batched step took 1.39 against 1.32 ms).
constexpr int kBatchItemMp = 2048;
__host__ __device__ constexpr int batch_item_cap(int S) { return S == kBsCols ? kBatchSmallMp : kBatchItemMp; }
__host__ __device__ constexpr int batch_item_lines(int S, int Mp) {
    return Mp <= kBatchSmallMp ? batch_item_cap(S) / Mp : 1;
}

struct BatchSched {
    int ctr[kBsStages * kBsClasses];    // next item of each pass and class (cleared by kb_sched every step)
    int total[kBsStages * kBsClasses];  // items of each pass and class this step
};

// The problem of workgroup / item x: the largest p with pref[p] <= x (pref non-decreasing, x < pref[n]).
__device__ __forceinline__ int batch_find(const int *__restrict__ pref, int nprob, int x) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Geometry (and twiddles when the FFT length changed) of every problem: workgroup p of 1024 threads.
__global__ __launch_bounds__(1024) void kb_geometry(const BatchProb *__restrict__ bp, int D, int p, double ipi, int mni,
                                                    int n1d_max, int mp_max, double lb) {
    const BatchProb &b = bp[blockIdx.x];
    geometry_block<float>(b.mmpart, b.cblk, b.Y, b.first, b.inv, (long)b.N * D, D, p, ipi, mni, n1d_max, mp_max, lb,
                          b.dg, b.tw);
}

// Line counts of the FFT passes per problem and class, their exclusive prefixes
// (pref[s * (nprob + 1) + p], s = pass + kBsStages * class), the totals, and cleared counters.  One
// workgroup of 1024 threads; problems in chunks of 1024; scans by wavefront shuffles, then over the
// wavefront totals.
__global__ __launch_bounds__(1024) void kb_sched(const BatchProb *__restrict__ bp, int nprob,
                                                 BatchSched *__restrict__ sch, int *__restrict__ pref) {
    constexpr int NS = kBsStages * kBsClasses;
    constexpr int NW = 1024 / kWave;
    __shared__ int wt[NS][NW];
    __shared__ int carry[NS];
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    if (t < NS) carry[t] = 0;
    for (int p0 = 0; p0 < nprob; p0 += 1024) {
        const int p = p0 + t;
        int cnt[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) cnt[s] = 0;
        if (p < nprob) {
            const DevGeo *dg = bp[p].dg;
            if (dg->live) {
                const int Mp = dg->g.Mp, n1d = dg->g.n1d, H = dg->g.H;
                auto items = [&](int S, int lines) { const int nl = batch_item_lines(S, Mp); return (lines + nl - 1) / nl; };
                const int o = kBsStages * batch_class(Mp);
                if (dg->kspec_stale) {
                    cnt[o + kBsKernRows] = items(kBsKernRows, n1d);
                    cnt[o + kBsKernCols] = items(kBsKernCols, H);
                }
                cnt[o + kBsRowsFwd] = items(kBsRowsFwd, n1d);
                cnt[o + kBsCols] = items(kBsCols, H);
                cnt[o + kBsRowsInv] = items(kBsRowsInv, n1d);
            }
        }
        int inc[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            int v = cnt[s];
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(v, o, kWave);
                if (lane >= o) v += y;
            }
            inc[s] = v;
        }
        __syncthreads();  // the previous chunk's wavefront totals and carry have been read
        if (lane == kWave - 1) {
#pragma unroll
            for (int s = 0; s < NS; s++) wt[s][w] = inc[s];
        }
        __syncthreads();
        if (p < nprob) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                int before = carry[s];
                for (int u = 0; u < w; u++) before += wt[s][u];
                pref[s * (nprob + 1) + p] = before + inc[s] - cnt[s];
            }
        }
        __syncthreads();
        if (t < NS) {
            int tot = 0;
            for (int u = 0; u < NW; u++) tot += wt[t][u];
            carry[t] += tot;
        }
    }
    __syncthreads();
    if (t < NS) {
        pref[t * (nprob + 1) + nprob] = carry[t];
        sch->total[t] = carry[t];
        sch->ctr[t] = 0;
    }
}

// Problems whose line prefixes a persistent pass keeps in LDS (larger batches search global memory).
constexpr int kBatchPfLds = 512;

// Workgroup loop of a persistent FFT pass: takes items of stage S until they run out and calls
// line(problem, item index, problem changed).  An item is batch_item_lines consecutive lines; the lines
// fill the buffer from its start and the twiddles sit after the class's largest buffer (MAXTOT).  Every call of `line` ends with a barrier, so the item
// slot and the staged twiddles are free again when the next item is taken.  The next item is claimed
// before the current one is worked, so the counter's round trip overlaps the line; the problem of an
// item is found in an LDS copy of the stage's prefixes.
template <int S, typename F>
__device__ __forceinline__ void batch_items(BatchSched *__restrict__ sch, const int *__restrict__ pref, int nprob,
                                            F &&line) {
    __shared__ int s_item;
    __shared__ int s_pf[kBatchPfLds];
    const int *pf = pref + S * (nprob + 1);
    const bool pf_lds = nprob + 1 <= kBatchPfLds;
    if (pf_lds)
        for (int q = threadIdx.x; q <= nprob; q += blockDim.x) s_pf[q] = pf[q];
    const int total = sch->total[S];
    if (threadIdx.x == 0) s_item = atomicAdd(&sch->ctr[S], 1);
    __syncthreads();
    int it = s_item;
    int cur = -1;
    while (it < total) {
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(&sch->ctr[S], 1);
        const int p = pf_lds ? batch_find(s_pf, nprob, it) : batch_find(pf, nprob, it);
        line(p, it - (pf_lds ? s_pf[p] : pf[p]), p != cur);
        cur = p;
        if (threadIdx.x == 0) s_item = next;  // every thread read s_item before the line's barriers
        __syncthreads();
        it = s_item;
    }
}

template <int NTHR, int MAXTOT, int CLS>
__global__ __launch_bounds__(NTHR) void kb_kern_rows(const BatchProb *__restrict__ bp, int nprob,
                                                     BatchSched *__restrict__ sch, const int *__restrict__ pref,
                                                     int kind_q, int kind_1, double df) {
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    const cx<float> *twl = nullptr;
    batch_items<kBsKernRows + kBsStages * CLS>(sch, pref, nprob, [&](int p, int a, bool fresh) {
        const BatchProb &b = bp[p];
        const int Mp = b.dg->g.Mp;
        if (fresh) twl = stage_twiddles(buf + lspan<true>(MAXTOT), b.tw, Mp);
        const int nl = batch_item_lines(kBsKernRows, Mp), a0 = a * nl;
        kern_rows_line<float, NTHR, MAXTOT, true>(b.dg, buf, twl, kind_q, kind_1, df, b.Trow, a0,
                                                  min(nl, b.dg->g.n1d - a0));
    });
}

template <int NTHR, int MAXTOT, int CLS>
__global__ __launch_bounds__(NTHR) void kb_kern_cols(const BatchProb *__restrict__ bp, int nprob,
                                                     BatchSched *__restrict__ sch, const int *__restrict__ pref) {
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    const cx<float> *twl = nullptr;
    batch_items<kBsKernCols + kBsStages * CLS>(sch, pref, nprob, [&](int p, int k, bool fresh) {
        const BatchProb &b = bp[p];
        const int Mp = b.dg->g.Mp;
        if (fresh) twl = stage_twiddles(buf + lspan<true>(MAXTOT), b.tw, Mp);
        const int nl = batch_item_lines(kBsKernCols, Mp), k0 = k * nl;
        kern_cols_line<float, NTHR, MAXTOT, true>(b.dg, buf, twl, b.Trow, 1, b.KQ, b.K1, b.KS, b.KT, k0,
                                                  min(nl, b.dg->g.H - k0));
    });
}

template <int NTHR, int MAXTOT, int CLS>
__global__ __launch_bounds__(NTHR) void kb_rows_fwd(const BatchProb *__restrict__ bp, int nprob,
                                                    BatchSched *__restrict__ sch, const int *__restrict__ pref) {
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    const cx<float> *twl = nullptr;
    batch_items<kBsRowsFwd + kBsStages * CLS>(sch, pref, nprob, [&](int p, int u, bool fresh) {
        const BatchProb &b = bp[p];
        const int Mp = b.dg->g.Mp;
        if (fresh) twl = stage_twiddles(buf + lspan<true>(MAXTOT), b.tw, Mp);
        const int nl = batch_item_lines(kBsRowsFwd, Mp), u0 = u * nl;
        rows_fwd_line<float, NTHR, MAXTOT, true>(b.dg, buf, twl, (const float *)nullptr, 4, 4, b.spec, b.fixgrid, u0,
                                                 min(nl, b.dg->g.n1d - u0));
    });
}

template <int NTHR, int MAXTOT, int CLS>
__global__ __launch_bounds__(NTHR) void kb_cols(const BatchProb *__restrict__ bp, int nprob,
                                                BatchSched *__restrict__ sch, const int *__restrict__ pref) {
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    __shared__ double shz[4 * kColZLanes / kWave];
    const cx<float> *twl = nullptr;
    batch_items<kBsCols + kBsStages * CLS>(sch, pref, nprob, [&](int p, int k, bool fresh) {
        const BatchProb &b = bp[p];
        const int Mp = b.dg->g.Mp;
        if (fresh) twl = stage_twiddles(buf + lspan<true>(MAXTOT), b.tw, Mp);
        const int nl = batch_item_lines(kBsCols, Mp), k0 = k * nl;
        cols_f32_line<NTHR, MAXTOT>(b.dg, buf, twl, b.spec, b.KQ, b.K1, b.KS, b.KT, b.ospec, b.zpart, shz, k0,
                                    min(nl, b.dg->g.H - k0));
    });
}

template <int NTHR, int MAXTOT, int CLS>
__global__ __launch_bounds__(NTHR) void kb_rows_inv(const BatchProb *__restrict__ bp, int nprob,
                                                    BatchSched *__restrict__ sch, const int *__restrict__ pref) {
    extern __shared__ unsigned char smem_raw[];
    cx<float> *buf = reinterpret_cast<cx<float> *>(smem_raw);
    const cx<float> *twl = nullptr;
    batch_items<kBsRowsInv + kBsStages * CLS>(sch, pref, nprob, [&](int p, int u, bool fresh) {
        const BatchProb &b = bp[p];
        const int Mp = b.dg->g.Mp;
        if (fresh) twl = stage_twiddles(buf + lspan<true>(MAXTOT), b.tw, Mp);
        const int nl = batch_item_lines(kBsRowsInv, Mp), u0 = u * nl;
        rows_inv_line<float, NTHR, MAXTOT, true>(b.dg, buf, twl, b.ospec, 3, reinterpret_cast<float *>(b.gout), 4, u0,
                                                 min(nl, b.dg->g.n1d - u0));
    });
}

// Final Z of every problem: workgroup p of kThreads (the reduction of the inverse row pass's first
// workgroup in a single context's step).
__global__ __launch_bounds__(kThreads) void kb_zfinal(const BatchProb *__restrict__ bp, int df_one) {
    const BatchProb &b = bp[blockIdx.x];
    if (!b.dg->live) return;
    __shared__ double sh[kThreads / kWave];
    zfinal_block(b.dg, b.zpart, df_one, (double)b.N, b.zbuf, sh);
}

// Attractive term: workgroups of problem p are [pref[p], pref[p + 1]), one per row block.
__global__ __launch_bounds__(kAttThreads) __attribute__((amdgpu_waves_per_eu(4))) void kb_attr(
    const BatchProb *__restrict__ bp, int nprob, const int *__restrict__ pref, float inv_df, int df_one, float mdf) {
    __shared__ AttrLds<2> L;
    const int p = batch_find(pref, nprob, blockIdx.x);
    const BatchProb &b = bp[p];
    if (!b.dg->live) return;
    attr_win_block<2>(b.wstart, b.wpure, b.wkey, b.wval, b.wbmax, b.Y, b.N, b.win_R, b.win_NW, (const int *)nullptr,
                      (b.N + b.win_R - 1) / b.win_R, inv_df, df_one, mdf,
                      b.sd->exag, b.pos, blockIdx.x - pref[p], L);
}

template <int P>
__global__ __launch_bounds__(kSpreadThreads) __attribute__((amdgpu_waves_per_eu(P <= 3 ? 4 : 1))) void kb_spread(
    const BatchProb *__restrict__ bp, int nprob, const int *__restrict__ pref) {
    __shared__ SpreadLds<P, 2> L;
    const int p = batch_find(pref, nprob, blockIdx.x);
    const BatchProb &b = bp[p];
    if (!b.dg->live) return;
    spread_block<P, 2>(b.Y, b.N, b.spread_pts, b.dg, b.fixgrid, blockIdx.x - pref[p], L);
}

template <int P>
__global__ __launch_bounds__(kThreads) void kb_interp(const BatchProb *__restrict__ bp, int nprob,
                                                      const int *__restrict__ pref) {
    __shared__ double sh[kThreads / kWave];
    const int p = batch_find(pref, nprob, blockIdx.x);
    const BatchProb &b = bp[p];
    if (!b.dg->live) return;
    interp_block<P, 2>(b.Y, b.pos, b.gout, b.dg, b.zbuf, b.sd, b.dC, b.uY, b.gains, b.part, b.N, blockIdx.x - pref[p],
                       sh);
}

// Mean of every problem's updated Y: workgroup p of kThreads.
__global__ __launch_bounds__(kThreads) void kb_mean(const BatchProb *__restrict__ bp) {
    const BatchProb &b = bp[blockIdx.x];
    if (!b.dg->live) return;
    __shared__ double sh[kThreads / kWave];
    mean_block(b.part, (b.N + kInterpPts - 1) / kInterpPts, 2, (double)b.N, b.mean, sh);
}

__global__ __launch_bounds__(kThreads) void kb_center(const BatchProb *__restrict__ bp, int nprob,
                                                      const int *__restrict__ pref) {
    const int p = batch_find(pref, nprob, blockIdx.x);
    const BatchProb &b = bp[p];
    if (!b.dg->live) return;
    center_block<float>(b.Y, (long)b.N * 2, 2, b.first, b.mean, b.mmpart, blockIdx.x - pref[p], b.cblk);
}

}  // namespace fitsne
