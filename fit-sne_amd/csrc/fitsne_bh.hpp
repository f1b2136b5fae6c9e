// fitsne_bh.hpp -- Barnes-Hut t-SNE gradient (nbody_algo = 1) in FP64: the reference's SPTree built on
// the host, its traversal and the edge forces on the GPU (gfx950, wave64).
//
// The reference builds an SPTree over Y every iteration (sptree.cpp:57-84, :189-277): a 2^D-ary tree of
// axis-aligned cells, one point per leaf, each node with an online center of mass updated in insertion
// order; exact duplicates stop at the leaf that holds their first copy.  computeNonEdgeForces
// (sptree.cpp:322-354) walks it depth first for every point and summarises a cell when
// max half-width / |y - com| < theta; computeEdgeForces (sptree.cpp:358-376) adds the attractive term
// over the CSR P, and dC = pos_f - neg_f / sum_Q (tsne.cpp:596-640).
//
// Here the tree is built on the host by the same insertion rules (its shape and centers of mass depend
// on the insertion order, so it is built in that order) and flattened in depth-first pre-order: the
// first child of an internal node is the next node, and skip[k] is the node after k's subtree, so the
// GPU walk needs no stack and visits cells in the reference's order.  One thread per point walks the
// tree; one thread per row sums its edges in CSR order; Z is summed over the points in index order.
// Contraction is off and sqrt / divisions are correctly rounded, so the result is the reference's to
// the bit.
#pragma once

#include <cfloat>
#include <vector>

#include "fitsne_kernels.hpp"

namespace fitsne {

constexpr int kBhMaxDims = 3;

// ---------------------------------------------------------------- host: the tree

struct BhTree {
    int D = 2, nc = 4;
    const double *Y = nullptr;
    std::vector<double> c, w, m;  // per node: cell center, half-width, center of mass (D each)
    std::vector<unsigned> cum, size, pt;
    std::vector<char> leaf;
    std::vector<int> first;  // first of the node's 2^D consecutive children, or -1

    int add(const double *center, const double *hw) {
        const int k = (int)cum.size();
        for (int d = 0; d < D; d++) {
            c.push_back(center[d]);
            w.push_back(hw[d]);
            m.push_back(0.0);
        }
        cum.push_back(0);
        size.push_back(0);
        pt.push_back(0);
        leaf.push_back(1);
        first.push_back(-1);
        return k;
    }

    // sptree.cpp:235-262: children i take the upper half of dimension d when bit d of i is 0
    void subdivide(int k) {
        double nc_[kBhMaxDims], nw[kBhMaxDims];
        int f = -1;
        for (int ch = 0; ch < nc; ch++) {
            unsigned div = 1;
            for (int d = 0; d < D; d++) {
                nw[d] = .5 * w[k * D + d];
                nc_[d] = ((unsigned)ch / div) % 2 == 1 ? c[k * D + d] - .5 * w[k * D + d] : c[k * D + d] + .5 * w[k * D + d];
                div *= 2;
            }
            const int id = add(nc_, nw);
            if (ch == 0) f = id;
        }
        first[k] = f;
        if (size[k] == 1) {
            for (int ch = 0; ch < nc; ch++)
                if (insert(f + ch, pt[k])) break;
        }
        size[k] = 0;
        leaf[k] = 0;
    }

    // sptree.cpp:189-232
    bool insert(int k, unsigned i) {
        const double *p = Y + (long)i * D;
        for (int d = 0; d < D; d++)
            if (c[k * D + d] - w[k * D + d] > p[d] || c[k * D + d] + w[k * D + d] < p[d]) return false;
        cum[k]++;
        const double mult1 = (double)(cum[k] - 1) / (double)cum[k], mult2 = 1.0 / (double)cum[k];
        for (int d = 0; d < D; d++) m[k * D + d] *= mult1;
        for (int d = 0; d < D; d++) m[k * D + d] += mult2 * p[d];
        if (leaf[k] && size[k] < 1) {
            pt[k] = i;
            size[k] = 1;
            return true;
        }
        if (size[k] == 1) {
            bool same = true;
            for (int d = 0; d < D; d++)
                if (p[d] != Y[(long)pt[k] * D + d]) {
                    same = false;
                    break;
                }
            if (same) return true;
        }
        if (leaf[k]) subdivide(k);
        for (int ch = 0; ch < nc; ch++)
            if (insert(first[k] + ch, i)) return true;
        return false;
    }

    // sptree.cpp:57-84 and :275-277
    void build(const double *Yin, int N, int Din) {
        D = Din;
        nc = 1 << D;
        Y = Yin;
        c.clear(), w.clear(), m.clear(), cum.clear(), size.clear(), pt.clear(), leaf.clear(), first.clear();
        double mean[kBhMaxDims] = {0, 0, 0}, mn[kBhMaxDims], mx[kBhMaxDims], hw[kBhMaxDims];
        for (int d = 0; d < D; d++) {
            mn[d] = DBL_MAX;
            mx[d] = -DBL_MAX;
        }
        for (long n = 0; n < N; n++)
            for (int d = 0; d < D; d++) {
                const double v = Y[n * D + d];
                mean[d] += v;
                if (v < mn[d]) mn[d] = v;
                if (v > mx[d]) mx[d] = v;
            }
        for (int d = 0; d < D; d++) mean[d] /= (double)N;
        for (int d = 0; d < D; d++) hw[d] = std::fmax(mx[d] - mean[d], mean[d] - mn[d]) + 1e-5;
        add(mean, hw);
        for (unsigned i = 0; i < (unsigned)N; i++) insert(0, i);
    }
};

// The walk's view of the tree in depth-first pre-order.
struct BhFlat {
    std::vector<double> com, maxw;
    std::vector<unsigned> cum;
    std::vector<int> leafpt, isleaf, skip;
    int depth = 0;

    void flatten(const BhTree &t) {
        const int n = (int)t.cum.size(), D = t.D;
        com.resize((size_t)n * D);
        maxw.resize(n);
        cum.resize(n);
        leafpt.resize(n);
        isleaf.resize(n);
        skip.resize(n);
        // iterative pre-order: (tree node, depth); a node's slot is taken when it is popped, and its skip
        // is filled when the walk has passed its subtree
        std::vector<std::pair<int, int>> st;
        std::vector<int> open;  // slots whose subtree is still being laid out, with their depth
        std::vector<int> open_depth;
        st.push_back({0, 0});
        int next = 0;
        depth = 0;
        while (!st.empty()) {
            const int k = st.back().first, dep = st.back().second;
            st.pop_back();
            while (!open.empty() && open_depth.back() >= dep) {
                skip[open.back()] = next;
                open.pop_back();
                open_depth.pop_back();
            }
            const int s = next++;
            for (int d = 0; d < D; d++) com[(size_t)s * D + d] = t.m[(size_t)k * D + d];
            double mw = 0.0;
            for (int d = 0; d < D; d++) {
                const double cur = t.w[(size_t)k * D + d];
                mw = mw > cur ? mw : cur;
            }
            maxw[s] = mw;
            cum[s] = t.cum[k];
            isleaf[s] = t.leaf[k] ? 1 : 0;
            leafpt[s] = (t.leaf[k] && t.size[k] == 1) ? (int)t.pt[k] : -1;
            if (dep + 1 > depth) depth = dep + 1;
            open.push_back(s);
            open_depth.push_back(dep);
            if (!t.leaf[k])
                for (int ch = t.nc - 1; ch >= 0; ch--) st.push_back({t.first[k] + ch, dep + 1});
        }
        while (!open.empty()) {
            skip[open.back()] = -1;
            open.pop_back();
            open_depth.pop_back();
        }
    }
};

// ---------------------------------------------------------------- device

// computeNonEdgeForces for every point (sptree.cpp:322-354): neg_f[i] (D) and Q[i].  The skip links
// replace the recursion; the order of the summarised cells is the reference's.  With neg_f null only Q.
template <int D>
__global__ __launch_bounds__(kThreads) void k_bh_nonedge(const double *__restrict__ Y, int N,
                                                         const double *__restrict__ com,
                                                         const double *__restrict__ maxw,
                                                         const unsigned *__restrict__ cum,
                                                         const int *__restrict__ leafpt,
                                                         const int *__restrict__ isleaf,
                                                         const int *__restrict__ skip, double theta,
                                                         double *__restrict__ neg_f, double *__restrict__ Q) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    double y[D], nf[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        y[d] = Y[(long)i * D + d];
        nf[d] = 0.0;
    }
    double q = 0.0;
    int k = 0;
    while (k >= 0) {
        if (cum[k] == 0u || leafpt[k] == i) {
            k = skip[k];
            continue;
        }
        double dd = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) dd += (y[d] - com[(long)k * D + d]) * (y[d] - com[(long)k * D + d]);
        if (isleaf[k] || maxw[k] / __dsqrt_rn(dd) < theta) {
            dd = 1.0 / (1.0 + dd);
            double mult = (double)cum[k] * dd;
            q += mult;
            mult *= dd;
#pragma unroll
            for (int d = 0; d < D; d++) nf[d] += mult * (y[d] - com[(long)k * D + d]);
            k = skip[k];
        } else {
            k = k + 1;  // first child
        }
    }
    Q[i] = q;
    if (neg_f) {
#pragma unroll
        for (int d = 0; d < D; d++) neg_f[(long)i * D + d] = nf[d];
    }
}

// computeEdgeForces (sptree.cpp:358-376): one thread per row, its entries in CSR order.
template <int D>
__global__ __launch_bounds__(kThreads) void k_bh_edge(const int64_t *__restrict__ rowP,
                                                      const uint32_t *__restrict__ colP,
                                                      const double *__restrict__ val, const double *__restrict__ Y,
                                                      int N, double *__restrict__ pos_f) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * kThreads + threadIdx.x;
    if (n >= N) return;
    double y[D], pf[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        y[d] = Y[(long)n * D + d];
        pf[d] = 0.0;
    }
    for (long e = rowP[n]; e < rowP[n + 1]; e++) {
        const long j = colP[e];
        double dd = 1.0;
#pragma unroll
        for (int d = 0; d < D; d++) dd += (y[d] - Y[j * D + d]) * (y[d] - Y[j * D + d]);
        dd = val[e] / dd;
#pragma unroll
        for (int d = 0; d < D; d++) pf[d] += dd * (y[d] - Y[j * D + d]);
    }
#pragma unroll
    for (int d = 0; d < D; d++) pos_f[(long)n * D + d] = pf[d];
}

// sum_Q = sum of Q[i] in index order (tsne.cpp:612-614), one thread; out[0] = sum_Q.
__global__ void k_seq_sum(const double *__restrict__ v, int n, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = .0;
    for (int i = 0; i < n; i++) s += v[i];
    out[0] = s;
    out[1] = 1.0 / s;
}

// dC = pos_f - neg_f / sum_Q (tsne.cpp:626-627).
__global__ __launch_bounds__(kThreads) void k_bh_combine(const double *__restrict__ pos_f,
                                                         const double *__restrict__ neg_f,
                                                         const double *__restrict__ zbuf, long n,
                                                         double *__restrict__ dC) {
    const double Z = zbuf[0];
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < n; e += (long)gridDim.x * kThreads)
        dC[e] = pos_f[e] - (neg_f[e] / Z);
}

}  // namespace fitsne
