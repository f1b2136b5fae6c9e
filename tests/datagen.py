"""Small seeded inputs for the CPU and GPU tests (Gaussian mixtures, CSR affinities, blob embeddings)."""
import numpy as np

from orc import knn_exact, normalize, rows_from_knn, symmetrize


def mixture(N, D, k, seed, sd=1.0, spread=4.0):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, spread, size=(k, D))
    labels = rng.integers(0, k, size=N)
    return centres[labels] + rng.normal(0.0, sd, size=(N, D)), labels


def affinities(X, perplexity=30.0, K=None):
    """kNN (exact) -> perplexity rows -> symmetrise -> normalise, as tsne.cpp:301-330 with knn_algo=2."""
    K = int(3 * perplexity) if K is None else K
    idx, dist = knn_exact(X, K)
    row, col, val = rows_from_knn(idx, dist, perplexity)
    return normalize(*symmetrize(row, col, val))


def blobs(N, D, k, seed, span=20.0, sd=1.5):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-span / 2, span / 2, size=(k, D))
    labels = rng.integers(0, k, size=N)
    Y = centres[labels] + rng.normal(0.0, sd, size=(N, D))
    return Y - Y.mean(axis=0)


def random_csr(N, per_row, seed, symmetric=True):
    """Random sparse affinities: per_row random neighbours per point, symmetrised and normalised."""
    rng = np.random.default_rng(seed)
    row = np.arange(N + 1, dtype=np.uint32) * per_row
    col = np.empty(N * per_row, dtype=np.uint32)
    for i in range(N):
        c = rng.choice(N - 1, size=per_row, replace=False)
        c[c >= i] += 1
        col[i * per_row:(i + 1) * per_row] = c
    val = rng.uniform(0.1, 1.0, size=N * per_row)
    if symmetric:
        return normalize(*symmetrize(row, col, val))
    return row, col, val / val.sum()


def dense_from_csr(P, N):
    row, col, val = P
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(row[i], row[i + 1]):
            A[i, col[k]] += val[k]
    return A
