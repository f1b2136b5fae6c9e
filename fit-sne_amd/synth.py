"""Synthetic workloads of the benchmark configurations (seeded, generated in place; nothing large ships).

c2_workload: N = 1,000,000 points in D = 50, 50 Gaussian clusters with Dirichlet(1) weights, centres
N(0, 6^2 I), sd 1.  Affinities: each point draws 45 random same-cluster neighbours, weights
exp(-d^2 / median_row(d^2)) normalised per row, then P = (P + P^T) / 2 normalised to sum 1 (about 90
non-zeros per row, int64 row offsets, int32 columns, float64 values).  Embedding: 2D blobs, one per
cluster, centres uniform in [-70, 70]^2, sd 3, zero mean (span about [-80, 80]).
"""
import sys
import time

import numpy as np


def _log(msg):
    print(f"[synth {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def symmetrize_normalize(N, rows, cols, vals):
    """(P + P^T) / 2 as CSR (duplicates summed), then val /= sum(val)."""
    import scipy.sparse as sp

    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    A = (A + A.T).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    val = A.data.astype(np.float64) * 0.5
    val /= val.sum()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), val


def cluster_affinities(X, labels, per_point, rng, chunk=1 << 20):
    N = X.shape[0]
    k = int(labels.max()) + 1
    members = [np.flatnonzero(labels == c) for c in range(k)]
    nbr = np.empty((N, per_point), dtype=np.int64)
    for c in range(k):
        idx = members[c]
        if idx.size < 2:
            nbr[idx] = idx[:, None]
            continue
        pick = rng.integers(0, idx.size - 1, size=(idx.size, per_point))
        self_pos = np.arange(idx.size)[:, None]
        pick = pick + (pick >= self_pos)  # skip self
        nbr[idx] = idx[pick]
    rows = np.repeat(np.arange(N, dtype=np.int64), per_point)
    cols = nbr.reshape(-1)
    d2 = np.empty(rows.size, dtype=np.float32)
    for s in range(0, rows.size, chunk):
        e = min(rows.size, s + chunk)
        diff = X[rows[s:e]] - X[cols[s:e]]
        d2[s:e] = np.einsum("ij,ij->i", diff, diff)
    d2 = d2.reshape(N, per_point).astype(np.float64)
    med = np.median(d2, axis=1, keepdims=True)
    w = np.exp(-d2 / np.maximum(med, 1e-12))
    w /= w.sum(axis=1, keepdims=True)
    return rows, cols, w.reshape(-1)


def pca_start(X, dims=2):
    """fast_tsne.py's PCA initialisation: the first `dims` principal components, scaled so that the first
    has standard deviation 1e-4 (the eigenvectors of the covariance, FP64)."""
    Xc = X.astype(np.float64)
    Xc -= Xc.mean(axis=0)
    w, v = np.linalg.eigh(Xc.T @ Xc)
    Y = Xc @ v[:, ::-1][:, :dims]
    Y /= np.std(Y[:, 0])
    return Y * 0.0001


def c2_workload(N=1_000_000, D=50, k=50, per_point=45, seed=20261204, pca=False):
    """pca=True: also return the PCA start of the D = 50 data (a fourth value)."""
    rng = np.random.default_rng(seed)
    t0 = time.time()
    weights = rng.dirichlet(np.ones(k))
    labels = rng.choice(k, size=N, p=weights)
    centres = rng.normal(0.0, 6.0, size=(k, D)).astype(np.float32)
    X = centres[labels] + rng.normal(0.0, 1.0, size=(N, D)).astype(np.float32)
    _log(f"C2 data {N}x{D} in {time.time() - t0:.1f}s")
    rows, cols, vals = cluster_affinities(X, labels, per_point, rng)
    Y0 = pca_start(X) if pca else None
    del X
    _log(f"C2 kNN-style affinities ({rows.size} directed edges) in {time.time() - t0:.1f}s")
    row_P, col_P, val_P = symmetrize_normalize(N, rows, cols, vals)
    _log(f"C2 symmetric P nnz={val_P.size} ({val_P.size / N:.1f}/row) in {time.time() - t0:.1f}s")
    c2 = rng.uniform(-70.0, 70.0, size=(k, 2))
    Y = c2[labels] + rng.normal(0.0, 3.0, size=(N, 2))
    Y -= Y.mean(axis=0)
    if pca:
        return (row_P, col_P, val_P), Y, labels, Y0
    return (row_P, col_P, val_P), Y, labels


def c3_workload(N=2_000_000, k=100, per_point=75, seed=20261205):
    """C3's gradient inputs: N points in 1D from k clusters with Dirichlet(0.5) weights, centres uniform
    in [-280, 280], sd 5 (span about [-300, 300]); P built like c2_workload's from per_point random
    same-cluster neighbours with weights on the cluster-local 1D gaps (int64 row offsets, int32
    columns), about 2 * per_point non-zeros per row: 75 neighbours give the ~150 per row that K = 150
    (perplexity 50) leaves after symmetrisation.  The kNN step of the full configuration (D = 100) is
    not part of the gradient path."""
    rng = np.random.default_rng(seed)
    t0 = time.time()
    weights = rng.dirichlet(np.full(k, 0.5))
    labels = rng.choice(k, size=N, p=weights)
    centres = rng.uniform(-280.0, 280.0, size=k)
    Y = centres[labels] + rng.normal(0.0, 5.0, size=N)
    X = (Y[:, None] + rng.normal(0.0, 1.0, size=(N, 1))).astype(np.float32)
    rows, cols, vals = cluster_affinities(X, labels, per_point, rng)
    _log(f"C3 affinities ({rows.size} directed edges) in {time.time() - t0:.1f}s")
    P = symmetrize_normalize(N, rows, cols, vals)
    _log(f"C3 symmetric P nnz={P[2].size} in {time.time() - t0:.1f}s")
    Y -= Y.mean()
    return P, Y.reshape(N, 1), labels


def c5_mixture(N=16_000, D=100, k=40, seed=1):
    """C5-style data: k equal-weight Gaussian clusters in D dimensions, centres N(0, 6^2 I), per-cluster
    sd log-uniform in [0.3, 3]."""
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, k, size=N)
    centres = rng.normal(0.0, 6.0, size=(k, D))
    sds = np.exp(rng.uniform(np.log(0.3), np.log(3.0), size=k))
    return centres[labels] + rng.normal(size=(N, D)) * sds[labels][:, None], labels


def c4_problems(n=64, nmin=2_000, nmax=200_000, per_point=30, seed=20261206):
    """C4: n independent embedding problems with N log-uniform in [nmin, nmax]; each a D = 50 mixture of
    5-30 equal-weight Gaussian clusters (centres N(0, 6^2 I), sd 1), P from per_point random
    same-cluster neighbours (as c2_workload), Y 2D blobs spanning about [-40, 40]."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n)).astype(int)
    out = []
    t0 = time.time()
    for i, N in enumerate(sizes):
        k = int(rng.integers(5, 31))
        labels = rng.integers(0, k, size=N)
        X = (rng.normal(0.0, 6.0, size=(k, 50))[labels] + rng.normal(0.0, 1.0, size=(N, 50))).astype(np.float32)
        rows, cols, vals = cluster_affinities(X, labels, per_point, rng)
        P = symmetrize_normalize(int(N), rows, cols, vals)
        cen = rng.uniform(-35.0, 35.0, size=(k, 2))
        Y = cen[labels] + rng.normal(0.0, 2.0, size=(N, 2))
        Y -= Y.mean(axis=0)
        out.append((P, Y))
    _log(f"C4: {n} problems, N in [{sizes.min()}, {sizes.max()}], total {sizes.sum()} points in {time.time() - t0:.1f}s")
    return out


def c1_data(N=10_000, D=50, k=10, seed=20261202):
    """C1: N points in D dimensions, k equal-weight isotropic Gaussians, centres N(0, 4^2 I), sd 1."""
    rng = np.random.default_rng(seed)
    labels = np.arange(N) % k
    rng.shuffle(labels)
    centres = rng.normal(0.0, 4.0, size=(k, D))
    X = centres[labels] + rng.normal(0.0, 1.0, size=(N, D))
    return X, labels


def small_mixture_workload(N=10_000, D=2, k=10, per_point=30, span=40.0, seed=7):
    """A scaled-down C2-like workload for quick checks (same construction)."""
    rng = np.random.default_rng(seed)
    weights = rng.dirichlet(np.ones(k))
    labels = rng.choice(k, size=N, p=weights)
    X = (rng.normal(0.0, 6.0, size=(k, 50))[labels] + rng.normal(0.0, 1.0, size=(N, 50))).astype(np.float32)
    rows, cols, vals = cluster_affinities(X, labels, per_point, rng)
    P = symmetrize_normalize(N, rows, cols, vals)
    cen = rng.uniform(-span / 2, span / 2, size=(k, D))
    Y = cen[labels] + rng.normal(0.0, span / 40, size=(N, D))
    Y -= Y.mean(axis=0)
    return P, Y, labels
