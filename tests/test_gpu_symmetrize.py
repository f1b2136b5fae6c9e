"""fitsne_symmetrize (symmetrizeMatrix, tsne.cpp:1730-1828, on the GPU) against the oracle.

The oracle's orc_symmetrize is pinned bit for bit to the reference's own symmetrizeMatrix
(tests/test_oracle_ref.py).  The GPU result must be identical: same row offsets, same column order in
every row, same values (the additions and the halving are the reference's operations).  Cases: kNN
graphs (every row K entries, mutual and one-sided neighbours), a ragged random CSR with empty rows and
self-loops, a single row, and a kNN graph of 100,000 points.
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import orc

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[symmetrize test] {msg}", flush=True)


def _knn_csr(N, K, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, 6)) + rng.normal(0, 3, size=(8, 6))[rng.integers(0, 8, N)]
    idx, dist = orc.knn_exact(X, K)
    row, col, val = orc.rows_from_knn(idx, dist, perplexity=K / 3.0)
    return row, col, val


def _ragged_csr(N, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, N)
    lens[rng.integers(0, N, N // 10)] = 0  # empty rows
    row = np.zeros(N + 1, dtype=np.uint32)
    row[1:] = np.cumsum(lens)
    cols = []
    for n in range(N):
        c = rng.choice(N, size=lens[n], replace=False)  # no repeated (row, col) entry
        if lens[n] and rng.random() < 0.05:
            c[0] = n  # a self-loop
        cols.append(c)
    col = np.concatenate(cols).astype(np.uint32) if cols else np.zeros(0, np.uint32)
    val = rng.random(col.size)
    return row, col, val


def _check(row, col, val):
    sr, sc, sv = fg.symmetrize(row, col, val)
    rr, rc, rv = orc.symmetrize(row, col, val)
    np.testing.assert_array_equal(sr, rr)
    np.testing.assert_array_equal(sc, rc)
    np.testing.assert_array_equal(sv, rv)
    return sr


@pytest.mark.parametrize("N,K,seed", [(3000, 30, 1), (1000, 90, 2), (200, 5, 3)])
def test_knn_graphs_match_the_oracle(N, K, seed):
    _note(f"kNN graph N={N} K={K}")
    row, col, val = _knn_csr(N, K, seed)
    sr = _check(row, col, val)
    assert N * K <= sr[-1] <= 2 * N * K


def test_ragged_csr_with_empty_rows_and_self_loops():
    _note("ragged CSR N=2500")
    _check(*_ragged_csr(2500, 4))


def test_single_row_and_empty_matrix():
    _note("N=1 self-loop; N=4 empty")
    _check(np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([0.25]))
    sr, sc, sv = fg.symmetrize(np.zeros(5, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    assert (sr == 0).all() and sc.size == 0 and sv.size == 0


def test_large_knn_graph_matches_the_oracle():
    N, K = 100_000, 45
    _note(f"kNN graph N={N} K={K}")
    rng = np.random.default_rng(5)
    # neighbour lists drawn inside clusters (as the benchmark workloads build them), distinct per row
    lab = rng.integers(0, 200, N)
    order = np.argsort(lab, kind="stable")
    pos = np.empty(N, np.int64)
    pos[order] = np.arange(N)
    start = np.searchsorted(lab[order], np.arange(200))
    size = np.bincount(lab, minlength=200)
    col = np.empty((N, K), np.uint32)
    for n in range(N):
        c = lab[n]
        pick = rng.choice(size[c] - 1, size=K, replace=False)
        pick = pick + (pick >= pos[n] - start[c])  # skip the point itself
        col[n] = order[start[c] + pick]
    row = (np.arange(N + 1) * K).astype(np.uint32)
    val = rng.random(N * K)
    _check(row, col.reshape(-1), val)
