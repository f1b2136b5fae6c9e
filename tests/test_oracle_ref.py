"""Pins oracle functions against the reference's own code (oracle/_ref, compiled from the reference
sources: the functions on the path that never reach FFTW).  Bit-exact where the arithmetic is the same;
these tests skip when the reference sources were not present at build time, and the committed golden
fixtures (tests/golden/) keep the same checks alive without them (test_golden.py)."""
import numpy as np
import pytest

import orc
from datagen import mixture, random_csr

pytestmark = pytest.mark.skipif(not orc.have_ref(), reason="oracle/_ref not built (reference sources absent)")


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_interpolate_bit_exact(p):
    rng = np.random.default_rng(p)
    y = np.concatenate([rng.uniform(0, 1, 500), [0.0, 1.0, -0.01, 1.02]])
    h = 1.0 / p
    sp = np.array([h / 2 + i * h for i in range(p)])
    # spacings accumulate like nbodyfft.cpp:31-34
    sp[0] = h / 2
    for i in range(1, p):
        sp[i] = sp[i - 1] + h
    a = orc.interpolate(p, y, sp, lib="oracle")
    b = orc.interpolate(p, y, sp, lib="ref")
    assert np.array_equal(a, b)


def test_kernels_bit_exact():
    rng = np.random.default_rng(1)
    for kind in (0, 1, 2):
        for df in (1.0, 0.5, 3.0):
            for _ in range(50):
                a, b, c, d = rng.normal(scale=5, size=4)
                assert orc.oracle().orc_kernel_1d(kind, a, b, df) == orc.ref().ref_kernel_1d(kind, a, b, df)
                assert orc.oracle().orc_kernel_2d(kind, a, b, c, d, df) == orc.ref().ref_kernel_2d(kind, a, b, c, d, df)


@pytest.mark.parametrize("perp,K,squared", [(30.0, 90, 0), (5.0, 15, 1), (50.0, 150, 0), (2.0, 7, 1)])
def test_distances2similarities_bit_exact(perp, K, squared):
    rng = np.random.default_rng(K)
    for trial in range(20):
        D = np.sort(rng.gamma(2.0, 1.0 + trial, size=K))
        P1 = np.empty(K)
        P2 = np.empty(K)
        b1 = orc.oracle().orc_distances2similarities(D.copy(), P1, K, -1, perp, -1.0, squared)
        b2 = orc.ref().ref_distances2similarities(D.copy(), P2, K, -1, perp, -1.0, squared)
        assert b1 == b2
        assert np.array_equal(P1, P2)


def test_distances2similarities_sigma_mode_and_list_bit_exact():
    rng = np.random.default_rng(9)
    D = np.sort(rng.gamma(2.0, 1.0, size=40))
    for sig in (0.5, 2.0):
        a, b = np.empty(40), np.empty(40)
        assert orc.oracle().orc_distances2similarities(D, a, 40, -1, -1.0, sig, 0) == \
            orc.ref().ref_distances2similarities(D.copy(), b, 40, -1, -1.0, sig, 0)
        assert np.array_equal(a, b)
    pl = np.array([5.0, 15.0, 10.0])
    a, b = np.empty(40), np.empty(40)
    ba = orc.oracle().orc_distances2similarities_list(D, a, 40, -1, 0.0, -1.0, 0, 3, pl)
    bb = orc.ref().ref_distances2similarities_list(D.copy(), b, 40, -1, 0.0, -1.0, 0, 3, pl.copy())
    assert ba == bb and np.array_equal(a, b)


def test_bisect_with_beta0_one_is_the_reference():
    rng = np.random.default_rng(4)
    D = np.sort(rng.gamma(2.0, 3.0, size=90))
    a, b = np.empty(90), np.empty(90)
    import ctypes as C
    it = C.c_int()
    H = C.c_double()
    ba = orc.oracle().orc_d2s_bisect(D, a, 90, 30.0, 1.0, 1, C.byref(it), C.byref(H))
    bb = orc.ref().ref_distances2similarities(D.copy(), b, 90, -1, 30.0, -1.0, 1)
    assert ba == bb and np.array_equal(a, b)
    assert abs(H.value - np.log(30.0)) < 1e-5


@pytest.mark.parametrize("N,k", [(50, 5), (300, 10), (1000, 12)])
def test_symmetrize_bit_exact(N, k):
    row, col, val = random_csr(N, k, seed=N, symmetric=False)
    a = orc.symmetrize(row, col, val, lib="oracle")
    b = orc.symmetrize(row, col, val, lib="ref")
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_vptree_perplexity_rows_match_exact_knn_rows():
    X, _ = mixture(400, 10, 4, seed=11)
    N, K, perp = 400, 30, 10.0
    row_r = np.empty(N + 1, dtype=np.uint32)
    col_r = np.empty(N * K, dtype=np.uint32)
    val_r = np.empty(N * K)
    orc.ref().ref_gaussian_perplexity_vptree(X.copy().reshape(-1), N, 10, perp, K, -1.0, 1, 0, np.zeros(1), 1,
                                             row_r, col_r, val_r)
    idx, dist = orc.knn_exact(X, K)
    row, col, val = orc.rows_from_knn(idx, dist, perp)
    assert np.array_equal(row, row_r)
    assert np.array_equal(col, col_r)
    assert np.array_equal(val, val_r)


def test_evaluate_error_fft_bit_exact():
    N = 500
    P = random_csr(N, 8, seed=5)
    rng = np.random.default_rng(6)
    Y = rng.normal(scale=3, size=(N, 2))
    for df in (1.0, 0.7):
        a = orc.kl(P, Y, 1234.5, df=df)
        b = orc.kl(P, Y, 1234.5, df=df, lib="ref")
        assert a == b
    Y1 = rng.normal(scale=3, size=(N, 1))
    assert orc.kl(P, Y1, 99.0) == orc.kl(P, Y1, 99.0, lib="ref")


def test_zero_mean_bit_exact():
    rng = np.random.default_rng(2)
    X = rng.normal(loc=3.0, size=(777, 3))
    a = X.copy().reshape(-1)
    b = X.copy().reshape(-1)
    orc.oracle().orc_zero_mean(a, 777, 3)
    orc.ref().ref_zero_mean(b, 777, 3)
    assert np.array_equal(a, b)
