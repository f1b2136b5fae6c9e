"""Pins of the oracle's Barnes-Hut restatement (nbody_algo = 1).

* Bit for bit against the reference's own computeGradient (tsne.cpp:596-640) over its SPTree
  (sptree.cpp) and its Barnes-Hut evaluateError (tsne.cpp:1359-1390), compiled from the reference's
  sources into oracle/_ref, with nthreads = 1: 1, 2 and 3 dimensions, theta 0.1 / 0.5 / 1.2, points on
  cell boundaries and exact duplicates (the reference's insertion order decides where a duplicate is
  counted).
* Against the mathematics, independently of the reference: as theta -> 0 every cell is opened down to
  its leaves, so the gradient tends to the exact O(N^2) gradient with the same P (dense); at theta = 1e-9
  they agree to rounding.
"""
import numpy as np
import pytest

import orc

needs_ref = pytest.mark.skipif(not orc.have_ref(), reason="reference sources not built into oracle/_ref")


def _csr(N, k, seed):
    rng = np.random.default_rng(seed)
    row = np.arange(N + 1, dtype=np.uint32) * k
    col = np.empty(N * k, dtype=np.uint32)
    for n in range(N):
        c = rng.choice(N - 1, size=k, replace=False)
        col[n * k:(n + 1) * k] = c + (c >= n)
    val = rng.random(N * k)
    return row, col, val / val.sum()


def _Y(N, D, seed, dup=False, grid=False):
    rng = np.random.default_rng(seed)
    Y = rng.normal(0, 4, size=(N, D))
    if grid:
        Y[: N // 4] = np.round(Y[: N // 4])  # many points on cell boundaries of a dyadic tree
    if dup:
        Y[10] = Y[3]
        Y[57] = Y[3]
        Y[200] = Y[150]
    return Y


@needs_ref
@pytest.mark.parametrize("D,theta,dup,grid", [(2, 0.5, False, False), (2, 0.1, True, False), (1, 0.5, True, True),
                                              (3, 1.2, False, True), (2, 0.5, True, True)])
def test_bh_gradient_and_error_bit_exact(D, theta, dup, grid):
    N = 600
    P = _csr(N, 7, 1)
    Y = _Y(N, D, 2, dup, grid)
    g, sq = orc.bh_gradient(P, Y, theta)
    gr, _ = orc.bh_gradient(P, Y, theta, lib="ref")
    np.testing.assert_array_equal(g, gr)
    assert sq > 0
    assert orc.bh_error(P, Y, theta) == orc.bh_error(P, Y, theta, lib="ref")


@pytest.mark.parametrize("D", [1, 2, 3])
def test_bh_tends_to_the_exact_gradient(D):
    N = 300
    row, col, val = _csr(N, 6, 3)
    Y = _Y(N, D, 4)
    dense = np.zeros((N, N))
    dense[np.repeat(np.arange(N), 6), col.astype(np.int64)] = val
    ge, ze = orc.exact_gradient(dense, Y)
    g, sq = orc.bh_gradient((row, col, val), Y, 1e-9)
    assert abs(sq - ze) <= 1e-12 * ze
    np.testing.assert_allclose(g, ge, rtol=0, atol=1e-12 * np.abs(ge).max())
    g5, _ = orc.bh_gradient((row, col, val), Y, 0.5)
    err = np.linalg.norm(g5 - ge) / np.linalg.norm(ge)
    assert 1e-8 < err < 0.3  # an approximation at theta = 0.5 (10% in 3D at this N), not the exact gradient
