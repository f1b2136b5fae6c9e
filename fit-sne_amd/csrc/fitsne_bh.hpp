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

// One cell: center and half-width (the reference's Cell corner / width), online center of mass,
// cum_size, at most one stored point, and the first of its 2^D consecutive children.
template <int D>
struct BhNode {
    double c[D], w[D], m[D];
    unsigned cum, pt;
    int first;
    unsigned char size, leaf;
};

template <int D>
struct BhTree {
    static constexpr int NC = 1 << D;
    const double *Y = nullptr;
    std::vector<BhNode<D>> nodes;

    int add(const double *center, const double *hw) {
        BhNode<D> n;
        for (int d = 0; d < D; d++) {
            n.c[d] = center[d];
            n.w[d] = hw[d];
            n.m[d] = 0.0;
        }
        n.cum = 0;
        n.pt = 0;
        n.first = -1;
        n.size = 0;
        n.leaf = 1;
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }

    // sptree.cpp:235-262: child ch takes the lower half of dimension d when bit d of ch is 1
    void subdivide(int k) {
        double nc[D], nw[D];
        int f = -1;
        for (int ch = 0; ch < NC; ch++) {
            unsigned div = 1;
            const BhNode<D> &p = nodes[k];
            for (int d = 0; d < D; d++) {
                nw[d] = .5 * p.w[d];
                nc[d] = ((unsigned)ch / div) % 2 == 1 ? p.c[d] - .5 * p.w[d] : p.c[d] + .5 * p.w[d];
                div *= 2;
            }
            const int id = add(nc, nw);  // may move nodes[]
            if (ch == 0) f = id;
        }
        nodes[k].first = f;
        if (nodes[k].size == 1) {
            const unsigned q = nodes[k].pt;
            for (int ch = 0; ch < NC; ch++)
                if (insert(f + ch, q)) break;
        }
        nodes[k].size = 0;
        nodes[k].leaf = 0;
    }

    // sptree.cpp:189-232
    bool insert(int k, unsigned i) {
        const double *p = Y + (size_t)i * D;
        {
            BhNode<D> &n = nodes[k];
            for (int d = 0; d < D; d++)
                if (n.c[d] - n.w[d] > p[d] || n.c[d] + n.w[d] < p[d]) return false;
            n.cum++;
            const double mult1 = (double)(n.cum - 1) / (double)n.cum, mult2 = 1.0 / (double)n.cum;
            for (int d = 0; d < D; d++) n.m[d] *= mult1;
            for (int d = 0; d < D; d++) n.m[d] += mult2 * p[d];
            if (n.leaf && n.size < 1) {
                n.pt = i;
                n.size = 1;
                return true;
            }
            if (n.size == 1) {
                bool same = true;
                for (int d = 0; d < D; d++)
                    if (p[d] != Y[(size_t)n.pt * D + d]) {
                        same = false;
                        break;
                    }
                if (same) return true;
            }
        }
        if (nodes[k].leaf) subdivide(k);
        const int f = nodes[k].first;
        for (int ch = 0; ch < NC; ch++)
            if (insert(f + ch, i)) return true;
        return false;
    }

    // sptree.cpp:57-84 and :275-277
    void build(const double *Yin, int N) {
        Y = Yin;
        nodes.clear();
        nodes.reserve((size_t)3 * N + 16);
        double mean[D], mn[D], mx[D], hw[D];
        for (int d = 0; d < D; d++) {
            mean[d] = 0.0;
            mn[d] = DBL_MAX;
            mx[d] = -DBL_MAX;
        }
        for (size_t n = 0; n < (size_t)N; n++)
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

// The walk's view of the tree in depth-first pre-order, written into one block of memory (the host
// side of a single copy): com[n*D] and maxw[n] (double), then cum[n], leafpt[n], isleaf[n], skip[n]
// (32-bit).  bh_flat_bytes gives its size for n nodes.
inline size_t bh_flat_bytes(size_t n, int D) { return n * (8 * (size_t)D + 8 + 16); }

struct BhFlatView {
    double *com, *maxw;
    unsigned *cum;
    int *leafpt, *isleaf, *skip;
    BhFlatView(void *base, size_t n, int D) {
        char *p = static_cast<char *>(base);
        com = reinterpret_cast<double *>(p);
        maxw = com + n * D;
        cum = reinterpret_cast<unsigned *>(maxw + n);
        leafpt = reinterpret_cast<int *>(cum + n);
        isleaf = leafpt + n;
        skip = isleaf + n;
    }
};

// Lay out node k's subtree from slot s in pre-order; returns the slot after it (the subtree's skip).
template <int D>
int bh_lay(const BhTree<D> &t, const BhFlatView &out, int k, int s) {
    const BhNode<D> &n = t.nodes[k];
    double mw = 0.0;
    for (int d = 0; d < D; d++) {
        out.com[(size_t)s * D + d] = n.m[d];
        mw = mw > n.w[d] ? mw : n.w[d];
    }
    out.maxw[s] = mw;
    out.cum[s] = n.cum;
    out.isleaf[s] = n.leaf ? 1 : 0;
    out.leafpt[s] = (n.leaf && n.size == 1) ? (int)n.pt : -1;
    int next = s + 1;
    if (!n.leaf)
        for (int ch = 0; ch < BhTree<D>::NC; ch++) next = bh_lay(t, out, n.first + ch, next);
    out.skip[s] = next;
    return next;
}

template <int D>
void bh_flatten(const BhTree<D> &t, const BhFlatView &out) {
    const int n = bh_lay(t, out, 0, 0);
    for (int s = 0; s < n; s++)
        if (out.skip[s] == n) out.skip[s] = -1;  // the end of the walk
}

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

// sum_Q = sum of Q[i] in index order (tsne.cpp:612-614): one workgroup stages chunks of v in LDS and
// its first thread adds them in order; out[0] = sum_Q.
constexpr int kSeqChunk = 8192;
__global__ __launch_bounds__(1024) void k_seq_sum(const double *__restrict__ v, int n, double *__restrict__ out) {
    __shared__ double buf[kSeqChunk];
    double s = .0;
    for (int base = 0; base < n; base += kSeqChunk) {
        const int len = min(kSeqChunk, n - base);
        __syncthreads();
        for (int i = threadIdx.x; i < len; i += blockDim.x) buf[i] = v[base + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < len; i++) s += buf[i];
    }
    if (threadIdx.x == 0) {
        out[0] = s;
        out[1] = 1.0 / s;
    }
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
