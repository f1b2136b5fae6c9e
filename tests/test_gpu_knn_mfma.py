"""The MFMA kNN (fitsne_knn_mfma: FP32 MFMA candidates, FP64 re-rank, FP64 kernel for unsettled queries)
against the FP64 kernel (fitsne_knn_exact, itself pinned to the oracle's VP-tree distances): the same
neighbours and distances, bit for bit, on C5-style mixtures (per-cluster sd log-uniform in [0.3, 3])
in D = 100 and D = 50, on tiny inputs, and with duplicated points (exact ties, broken by index)."""
import numpy as np
import pytest

import fitsne_gpu as fg
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("N,D,K,k", [(20000, 100, 90, 40), (12000, 50, 90, 20), (3000, 7, 30, 5)])
def test_mfma_knn_equals_the_fp64_kernel(N, D, K, k):
    X, _ = synth.c5_mixture(N=N, D=D, k=k, seed=3 + D)
    say(f"[gpu] knn_mfma N={N} D={D} K={K}")
    i_m, d_m, redo = fg.knn_mfma(X, K)
    i_e, d_e = fg.knn_exact(X, K)
    say(f"    recomputed by the FP64 kernel: {redo} of {N}")
    assert np.array_equal(i_m, i_e)
    assert np.array_equal(d_m, d_e)
    assert redo <= N // 20


def test_mfma_knn_with_duplicates_and_tiny_n():
    rng = np.random.default_rng(5)
    X = rng.normal(size=(400, 12))
    X[200:260] = X[0:60]  # exact duplicates: distance-0 ties broken by the smaller index
    for K in (1, 20, 100):
        i_m, d_m, _ = fg.knn_mfma(X, K)
        i_e, d_e = fg.knn_exact(X, K)
        assert np.array_equal(i_m, i_e) and np.array_equal(d_m, d_e)
    Xs = rng.normal(size=(40, 3))
    i_m, d_m, _ = fg.knn_mfma(Xs, 39)
    i_e, d_e = fg.knn_exact(Xs, 39)
    assert np.array_equal(i_m, i_e) and np.array_equal(d_m, d_e)
