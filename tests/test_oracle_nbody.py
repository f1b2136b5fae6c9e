"""Pins the oracle's n-body FFT path and gradients by what the mathematics fixes:
 - the circulant FFT step equals the direct linear convolution on the node lattice;
 - spread/interpolate use the reference's Lagrange basis (pinned bit-exact in test_oracle_ref);
 - end to end, potentials and gradients sit within interpolation error of the exact O(N^2) sums,
   using the reference's own computeExactGradient (oracle/_ref) where available.
"""
import numpy as np
import pytest
from scipy.signal import convolve2d

import orc
from datagen import blobs, dense_from_csr, random_csr


def _weights(t, sp):
    return orc.interpolate(len(sp), t, sp)


def _numpy_nbody_2d(xs, ys, charges, nb, p, geo, kernel, df):
    """Spread -> direct linear convolution -> interpolate, independent of any FFT."""
    lo, hi, sp = geo["lo"], geo["hi"], geo["sp"]
    bw = hi[0] - lo[0]
    cmin = lo[0]
    N = xs.size
    n1d = nb * p
    xi = np.clip(((xs - cmin) / bw).astype(int), 0, nb - 1)
    yi = np.clip(((ys - cmin) / bw).astype(int), 0, nb - 1)
    tx = (xs - lo[yi * nb + xi]) / bw
    ty = (ys - lo[nb * nb + yi * nb + xi]) / bw
    wx, wy = _weights(tx, sp), _weights(ty, sp)
    nt = charges.shape[1]
    W = np.zeros((n1d, n1d, nt))
    for a in range(p):
        for b in range(p):
            np.add.at(W, (xi * p + a, yi * p + b), (wx[a] * wy[b])[:, None] * charges)
    h = bw / p
    off = np.arange(-(n1d - 1), n1d) * h
    Kmat = np.array([[orc.oracle().orc_kernel_2d(kernel, 0.0, 0.0, u, v, df) for v in off] for u in off])
    V = np.stack([convolve2d(W[:, :, d], Kmat, mode="full")[n1d - 1:2 * n1d - 1, n1d - 1:2 * n1d - 1]
                  for d in range(nt)], axis=-1)
    pot = np.zeros((N, nt))
    for a in range(p):
        for b in range(p):
            pot += (wx[a] * wy[b])[:, None] * V[xi * p + a, yi * p + b]
    return pot


@pytest.mark.parametrize("kernel,df", [(orc.SQUARED_CAUCHY, 1.0), (orc.GENERAL, 0.6), (orc.SQUARED_GENERAL, 2.0)])
def test_n_body_fft_2d_equals_direct_convolution(kernel, df):
    rng = np.random.default_rng(kernel)
    N, nb, p = 400, 25, 3
    Y = rng.uniform(-10, 10, size=(N, 2))
    mn, mx = orc.minmax_quirk(Y)
    geo = orc.precompute_2d(mx, mn, nb, p, kernel, df)
    xs, ys = Y[:, 0].copy(), Y[:, 1].copy()
    ch = np.stack([np.ones(N), xs, ys, xs * xs + ys * ys], axis=1)
    got = orc.n_body_fft_2d(xs, ys, ch, nb, p, geo)
    want = _numpy_nbody_2d(xs, ys, ch, nb, p, geo, kernel, df)
    scale = np.abs(want).max(axis=0)
    assert np.all(np.abs(got - want).max(axis=0) <= 1e-12 * scale)


def test_precompute_2d_geometry_closed_form():
    geo = orc.precompute_2d(7.5, -2.5, 25, 3, orc.SQUARED_CAUCHY)
    bw = 10.0 / 25
    j = np.arange(25)
    assert np.allclose(geo["lo"][:25], j * bw - 2.5, atol=1e-14)
    assert np.allclose(geo["lo"][625:625 + 25], -2.5, atol=1e-14)  # y-lower of row 0
    assert np.allclose(geo["sp"], [1 / 6, 0.5, 5 / 6], atol=1e-15)
    assert np.allclose(geo["xt"], -2.5 + (np.arange(75) + 0.5) * bw / 3, atol=1e-12)
    # kernel spectrum is real up to rounding times (-1)^(k0+k1) (kernel centred at n1d, even)
    kt = geo["kt"]
    assert np.abs(kt.imag).max() <= 1e-10 * np.abs(kt.real).max()


def test_n_body_fft_1d_equals_direct_convolution():
    rng = np.random.default_rng(3)
    N, nb, p = 300, 40, 3
    Y = rng.uniform(-15, 25, size=N)
    for kernel, df in [(orc.SQUARED_CAUCHY, 1.0), (orc.GENERAL, 0.5), (orc.SQUARED_GENERAL, 0.5)]:
        geo = orc.precompute_1d(Y.min(), Y.max(), nb, p, kernel, df)
        ch = np.stack([np.ones(N), Y, Y * Y], axis=1)
        got = orc.n_body_fft_1d(Y, ch, nb, p, geo)
        lo, hi, sp = geo["lo"], geo["hi"], geo["sp"]
        bw = hi[0] - lo[0]
        bi = np.minimum(((Y - lo[0]) / bw).astype(int), nb - 1)
        w = _weights((Y - lo[bi]) / bw, sp)
        n1d = nb * p
        W = np.zeros((n1d, 3))
        for a in range(p):
            np.add.at(W, bi * p + a, w[a][:, None] * ch)
        h = bw / p
        Kd = np.array([[orc.oracle().orc_kernel_1d(kernel, 0.0, (r - c) * h, df) for c in range(n1d)] for r in range(n1d)])
        V = Kd @ W
        want = sum(w[a][:, None] * V[bi * p + a] for a in range(p))
        assert np.all(np.abs(got - want).max(axis=0) <= 1e-12 * np.abs(want).max(axis=0))


def test_minmax_else_if_quirk():
    Y = np.array([[-5.0, 1.0], [2.0, -3.0], [4.0, 0.5]])
    mn, mx = orc.minmax_quirk(Y)
    assert mx == 4.0 and mn == -3.0   # xs[0] = -5 set the first running max and is never a min
    Y2 = np.array([[0.0, 1.0], [2.0, -3.0]])
    assert orc.minmax_quirk(Y2) == (-3.0, 2.0)


def test_n_boxes_rounding_table():
    L = orc.oracle()
    assert L.orc_n_boxes_2d(-80.0, 80.0, 1.0, 50) == 175
    assert L.orc_n_boxes_2d(-5.0, 5.0, 1.0, 50) == 50
    assert L.orc_n_boxes_2d(-25.6, 25.6, 1.0, 50) == 55
    assert L.orc_n_boxes_2d(-120.0, 120.0, 1.0, 50) == 240
    assert L.orc_n_boxes_2d(-100.0, 100.0, 1.0, 50) == 200
    assert L.orc_n_boxes_1d(-300.0, 300.0, 1.0, 50) == 600
    assert L.orc_n_boxes_1d(-10.0, 10.7, 1.0, 50) == 50


def _exact_grad(P, Y, df):
    N = Y.shape[0]
    pos = orc.attractive(P, Y, df)
    neg, Z = orc.exact_repulsion(Y, df)
    return pos - neg, Z


@pytest.mark.parametrize("D,df", [(2, 1.0), (2, 0.5), (1, 1.0), (1, 0.5)])
def test_fft_gradient_within_interpolation_error_of_exact(D, df):
    N = 1500
    P = random_csr(N, 10, seed=D * 10 + int(df * 10))
    Y = blobs(N, D, 6, seed=4, span=30.0, sd=2.0)
    g, Z = orc.gradient(P, Y, df=df)
    ge, Ze = _exact_grad(P, Y, df)
    # Z: the reference subtracts N for the interpolated self-terms.  Measured interpolation error at
    # p=3, one box per unit: Z ~5e-5, gradient ~2e-3 (2D) and ~1e-2 (1D); a dropped term, a wrong sign
    # or a swapped axis gives O(1).
    assert abs(Z - Ze) / Ze < 2e-4
    rel = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    assert rel < (5e-3 if D == 2 else 2e-2), rel


def test_exact_gradient_formula_matches_reference():
    if not orc.have_ref():
        pytest.skip("oracle/_ref not built")
    N = 300
    P = random_csr(N, 6, seed=1)
    Y = blobs(N, 2, 3, seed=2, span=10.0)
    Pd = dense_from_csr(P, N)
    dC = np.empty(N * 2)
    orc.ref().ref_exact_gradient(Pd.reshape(-1).copy(), Y.reshape(-1).copy(), N, 2, dC, 1.0)
    ge, _ = _exact_grad(P, Y, 1.0)
    assert np.allclose(dC.reshape(N, 2), ge, rtol=0, atol=1e-12 * np.abs(ge).max())


def test_attractive_brute_force():
    N = 200
    P = random_csr(N, 7, seed=3)
    Y = blobs(N, 2, 3, seed=1)
    A = dense_from_csr(P, N)
    for df in (1.0, 0.3):
        d = Y[:, None, :] - Y[None, :, :]
        q = 1.0 / (1.0 + (d ** 2).sum(-1) / df)
        want = ((A * q)[:, :, None] * d).sum(1)
        assert np.allclose(orc.attractive(P, Y, df), want, rtol=0, atol=1e-14)


def test_update_closed_form():
    rng = np.random.default_rng(0)
    N, D = 100, 2
    Y = rng.normal(size=(N, D)).reshape(-1)
    uY = rng.normal(size=N * D)
    uY[:5] = 0.0
    gains = rng.uniform(0.005, 2, N * D)
    dY = rng.normal(size=N * D) * 10
    dY[5:8] = 0.0
    Y0, u0, g0 = Y.copy(), uY.copy(), gains.copy()
    orc.oracle().orc_update(Y, uY, gains, dY, N, D, 0.5, 200.0, 5.0, 0)
    sg = lambda x: np.sign(x)
    g1 = np.where(sg(dY) != sg(u0), g0 + 0.2, g0 * 0.8)
    g1 = np.maximum(g1, 0.01)
    u1 = 0.5 * u0 - 200.0 * g1 * dY
    u1 = u1.reshape(N, D)
    n = np.sqrt((u1 ** 2).sum(1))
    u1 = np.where((n > 5.0)[:, None], u1 * (5.0 / n)[:, None], u1).reshape(-1)
    assert np.allclose(gains, g1, rtol=0, atol=0)
    assert np.allclose(uY, u1, rtol=1e-15, atol=0)
    assert np.allclose(Y, Y0 + u1, rtol=1e-15, atol=1e-15)
    Y2 = Y0.copy()
    orc.oracle().orc_update(Y2, u0.copy(), g0.copy(), dY, N, D, 0.5, 200.0, 5.0, 1)
    assert np.array_equal(Y2, Y0 - dY)


def test_run_fft_small_embedding_separates_clusters():
    from datagen import mixture, affinities
    X, lab = mixture(600, 10, 3, seed=5, spread=6.0)
    P = affinities(X, perplexity=20.0)
    rng = np.random.default_rng(1)
    Y0 = rng.normal(scale=1e-4, size=(600, 2))
    Y, costs = orc.run_fft(P, Y0, max_iter=300, learning_rate=200.0)
    kls = costs[49::50]
    assert np.all(np.isfinite(kls)) and kls[-1] < kls[0]
    # clusters separate: nearest-centroid assignment recovers the labels
    cent = np.array([Y[lab == c].mean(0) for c in range(3)])
    assign = np.argmin(((Y[:, None, :] - cent[None]) ** 2).sum(-1), axis=1)
    assert np.mean(assign == lab) > 0.98
