"""Pins of the oracle's exact (theta = 0) restatement.

* Against the reference's own functions, compiled from its sources into oracle/_ref (bit for bit):
  computeSquaredEuclideanDistance (tsne.cpp:1832-1848), the dense computeGaussianPerplexity
  (tsne.cpp:1503-1530) in perplexity, perplexity-list and fixed-sigma modes, computeExactGradient
  (tsne.cpp:1232-1282) and evaluateError (tsne.cpp:1286-1326) for df = 1 and df != 1, D = 1, 2, 3.
* Against the mathematics, independently of the reference:
  - the symmetrisation of tsne.cpp:196-216 gives a symmetric matrix of total 1 equal to (R + R^T) / sum
    off the diagonal with R's diagonal / sum on it, checked by a brute-force loop;
  - the exact gradient is a quarter of the derivative of the exact KL: 4 dC = dKL/dY for symmetric P
    of total 1 (central finite differences), for df = 1 and df != 1.  A dropped term, a wrong sign or a
    transposed index in either function breaks this identity.
"""
import numpy as np
import pytest

import orc

needs_ref = pytest.mark.skipif(not orc.have_ref(), reason="reference sources not built into oracle/_ref")


def _data(N=120, D=9, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 3, size=(4, D))
    X = c[rng.integers(0, 4, N)] + rng.normal(size=(N, D))
    X -= X.mean(0)
    return X / np.abs(X).max()


def _embedding(N, D, seed, scale=3.0):
    return np.random.default_rng(seed).normal(0, scale, size=(N, D))


@needs_ref
def test_squared_distances_bit_exact():
    X = _data(77, 13, 1)
    X[5] = X[40]  # a duplicate: zero off-diagonal distance
    np.testing.assert_array_equal(orc.squared_distances(X), orc.squared_distances(X, lib="ref"))


@needs_ref
@pytest.mark.parametrize("mode", ["perplexity", "list", "sigma"])
def test_dense_gaussian_perplexity_bit_exact(mode):
    X = _data(150, 10, 2)
    kw = {"perplexity": dict(perplexity=20.0), "list": dict(perplexity=0.0, plist=[5.0, 25.0, 12.5]),
          "sigma": dict(perplexity=-1.0, sigma=0.3)}[mode]
    ours = orc.gaussian_perplexity_dense(X, **kw)
    theirs = orc.gaussian_perplexity_dense(X, lib="ref", **kw)
    np.testing.assert_array_equal(ours, theirs)
    # the self entry is DBL_MIN / sum_P, never exp(0)
    assert np.all(np.diag(ours) < 1e-290)


def test_exact_symmetrisation_brute_force():
    R = np.random.default_rng(3).random((40, 40))
    P, s = orc.exact_symmetrize_normalize(R)
    N = R.shape[0]
    S = R.copy()
    for n in range(N):
        for m in range(n + 1, N):
            S[n, m] = S[m, n] = R[n, m] + R[m, n]
    tot = 0.0
    for v in S.reshape(-1):
        tot += v
    assert s == tot
    np.testing.assert_array_equal(P, S / tot)
    np.testing.assert_array_equal(P, P.T)
    assert abs(P.sum() - 1.0) < 1e-13


@needs_ref
@pytest.mark.parametrize("D,df", [(2, 1.0), (1, 0.5), (3, 1.0), (2, 0.7)])
def test_exact_gradient_and_kl_bit_exact(D, df):
    X = _data(200, 8, 4)
    P, _ = orc.exact_P(X, 15.0)
    Y = _embedding(200, D, 5)
    g, sq = orc.exact_gradient(P, Y, df)
    gr, _ = orc.exact_gradient(P, Y, df, lib="ref")
    np.testing.assert_array_equal(g, gr)
    assert sq > 0
    assert orc.exact_kl(P, Y, df) == orc.exact_kl(P, Y, df, lib="ref")


@pytest.mark.parametrize("D,df", [(2, 1.0), (2, 0.6), (1, 1.0), (3, 1.5)])
def test_exact_gradient_is_quarter_kl_derivative(D, df):
    N = 60
    X = _data(N, 6, 6)
    P, _ = orc.exact_P(X, 10.0)
    Y = _embedding(N, D, 7, scale=1.5)
    g, _ = orc.exact_gradient(P, Y, df)
    h = 1e-6
    rng = np.random.default_rng(8)
    for _ in range(12):
        n, d = int(rng.integers(N)), int(rng.integers(D))
        Yp, Ym = Y.copy(), Y.copy()
        Yp[n, d] += h
        Ym[n, d] -= h
        fd = (orc.exact_kl(P, Yp, df) - orc.exact_kl(P, Ym, df)) / (2 * h)
        assert abs(4 * g[n, d] - fd) <= 1e-6 * max(1.0, abs(fd)), (n, d, 4 * g[n, d], fd)
