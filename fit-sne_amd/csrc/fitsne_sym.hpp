// fitsne_sym.hpp -- symmetrizeMatrix (tsne.cpp:1730-1828) on the GPU, with the reference's output
// order and values (gfx950, wave64).
//
// The reference visits the entries e = (n, c) in CSR order.  An entry whose partner (c, n) exists is
// placed only from the lower row (n <= c); an entry without a partner is always placed.  Placing e
// appends one output to row n and, when c != n, one to row c, with the value v[e] + v[partner] (or
// v[e]) halved at the end.  Row r of the output therefore lists the placed entries that touch r in
// increasing e: it is the sequence of keys (r, e) in sorted order.  The GPU form:
//   1. sort the keys (row, col) of all entries (stable radix sort, entry index as payload): each row's
//      entries stay in its own segment, ordered by column;
//   2. per entry, binary-search its partner in row c's segment (the last match, as the reference's
//      scan keeps the last one), mark it placed, emit its output keys (n, e) and (c, e) and count the
//      outputs per row (integer atomics: the counts are exact);
//   3. sort the output keys (stable radix sort; absent outputs carry a key past every row);
//   4. row offsets by an exclusive scan of the counts; output i = key (r, e): column = the other end of
//      e, value = (v[e] + v[partner]) / 2.
// Deterministic: every position follows from the sorted keys.
#pragma once

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "fitsne_kernels.hpp"

namespace fitsne {

// erow[e] = n for every entry e of row n.  One wavefront per row.
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_entry_rows(const TI *__restrict__ rowP, int N, int *__restrict__ erow) {
    const long r = ((long)blockIdx.x * kThreads + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (r >= N) return;
    for (long e = (long)rowP[r] + lane; e < (long)rowP[r + 1]; e += kWave) erow[e] = (int)r;
}

// keys[e] = n << cbits | col[e], vals[e] = e.
__global__ __launch_bounds__(kThreads) void k_sym_keys(const int *__restrict__ erow, const uint32_t *__restrict__ colP,
                                                       long nnz, int cbits, uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals) {
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < nnz; e += (long)gridDim.x * kThreads) {
        keys[e] = ((uint64_t)erow[e] << cbits) | colP[e];
        vals[e] = (uint32_t)e;
    }
}

// Partner search, placement and output keys (steps 2 of the header).  skeys/svals: the sorted
// (row, col) keys with their entry indices.  okeys[2e], okeys[2e + 1]: the outputs of e in row n and
// row c, or `none` (a key past every row).  cnt[r] (zeroed by the caller) counts the outputs of row r.
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_sym_place(const TI *__restrict__ rowP, const int *__restrict__ erow,
                                                        const uint32_t *__restrict__ colP,
                                                        const uint64_t *__restrict__ skeys,
                                                        const uint32_t *__restrict__ svals, long nnz, int cbits,
                                                        int ebits, uint64_t none, int64_t *__restrict__ partner,
                                                        int *__restrict__ cnt, uint64_t *__restrict__ okeys) {
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < nnz; e += (long)gridDim.x * kThreads) {
        const uint32_t n = (uint32_t)erow[e], c = colP[e];
        const uint64_t target = ((uint64_t)c << cbits) | n;
        // first position in row c's segment whose key exceeds target
        long lo = (long)rowP[c], hi = (long)rowP[c + 1];
        const long seg0 = lo;
        while (lo < hi) {
            const long mid = lo + (hi - lo) / 2;
            if (skeys[mid] <= target) lo = mid + 1;
            else hi = mid;
        }
        const int64_t m = (lo > seg0 && skeys[lo - 1] == target) ? (int64_t)svals[lo - 1] : (int64_t)-1;
        partner[e] = m;
        const bool placed = m < 0 || n <= c;
        uint64_t k0 = none, k1 = none;
        if (placed) {
            k0 = ((uint64_t)n << ebits) | (uint64_t)e;
            atomicAdd(&cnt[n], 1);
            if (c != n) {
                k1 = ((uint64_t)c << ebits) | (uint64_t)e;
                atomicAdd(&cnt[c], 1);
            }
        }
        okeys[2 * e] = k0;
        okeys[2 * e + 1] = k1;
    }
}

// Output i (< total) from its sorted key (r, e): the other end of e and (v[e] + v[partner]) / 2
// (tsne.cpp:1793-1794, :1801-1802, :1814).
__global__ __launch_bounds__(kThreads) void k_sym_fill(const uint64_t *__restrict__ okeys, const int *__restrict__ total,
                                                       int ebits, const int *__restrict__ erow,
                                                       const uint32_t *__restrict__ colP, const double *__restrict__ val,
                                                       const int64_t *__restrict__ partner,
                                                       uint32_t *__restrict__ scol, double *__restrict__ sval) {
    const long M = total[0];
    const uint64_t emask = (ebits >= 64) ? ~0ull : ((1ull << ebits) - 1ull);
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < M; i += (long)gridDim.x * kThreads) {
        const uint64_t k = okeys[i];
        const uint32_t r = (uint32_t)(k >> ebits);
        const long e = (long)(k & emask);
        const uint32_t n = (uint32_t)erow[e], c = colP[e];
        scol[i] = r == n ? c : n;
        const int64_t m = partner[e];
        const double v = m >= 0 ? val[e] + val[m] : val[e];
        sval[i] = v / 2.0;
    }
}

}  // namespace fitsne
