"""Barnes-Hut t-SNE (nbody_algo = 1) on the GPU against the oracle.

The oracle's SPTree restatement is bit-exact against the reference's own computeGradient and
evaluateError (tests/test_oracle_bh.py, through oracle/_ref).  The GPU path builds the same tree in the
same insertion order and walks it in the same order with the same rounding, so:
* the gradient equals the oracle's bit for bit (1, 2 and 3 dimensions, theta 0.2 / 0.5 / 1.0, points on
  cell boundaries, exact duplicates), and so does Z;
* the KL is within 1e-13 (Z summed per point, then over the points; the reference keeps one running sum);
* 20 optimisation steps (exaggeration, momentum, clip) stay within 1e-12 of the oracle's loop (the
  zeroMean sums differ in order);
* the drop-in program with nbody_algo = 'Barnes-Hut' runs end to end and its KL follows the oracle's.
"""
import math
import os

import numpy as np
import pytest

import fast_tsne as ft
import fitsne_gpu as fg
import orc

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[bh test] {msg}", flush=True)


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
        Y[: N // 4] = np.round(Y[: N // 4])
    if dup:
        Y[10] = Y[3]
        Y[57] = Y[3]
        Y[200] = Y[150]
    return Y


def _ctx(D, theta):
    c = fg.Context(no_dims=D, precision=fg.FP64_BH)
    c.set_theta(theta)
    return c


@pytest.mark.parametrize("N,D,theta,dup,grid", [(3000, 2, 0.5, False, False), (2500, 1, 0.2, True, True),
                                                (2000, 3, 1.0, True, False), (1500, 2, 0.5, True, True)])
def test_bh_gradient_bitwise_equals_the_oracle(N, D, theta, dup, grid):
    P = _csr(N, 9, 1)
    Y = _Y(N, D, 2, dup, grid)
    _note(f"gradient N={N} D={D} theta={theta} dup={dup} grid={grid}")
    c = _ctx(D, theta)
    c.set_P(*P)
    c.set_Y(Y)
    g, z = c.gradient()
    kl = c.kl()
    c.close()
    gr, zr = orc.bh_gradient(P, Y, theta)
    np.testing.assert_array_equal(g, gr)
    assert z == zr
    klr = orc.bh_error(P, Y, theta)
    assert abs(kl - klr) <= 1e-13 * abs(klr)


def test_bh_steps_match_the_oracle():
    N, D, theta = 1200, 2, 0.5
    P = _csr(N, 8, 3)
    Y0 = _Y(N, D, 4) * 1e-2
    early, stop, mom_switch, lr, msn, iters = 12.0, 8, 12, 200.0, 5.0, 20
    _note(f"steps N={N} iters={iters}")
    c = _ctx(D, theta)
    c.set_P(*P)
    c.set_exaggeration(early)
    c.set_Y(Y0)
    mom = 0.5
    for it in range(iters):
        c.step(mom, lr, msn)
        if it == stop:
            c.set_exaggeration(1.0)
        if it == mom_switch:
            mom = 0.8
    Yg = c.get_Y()
    c.close()
    L = orc.oracle()
    row, col, val = P
    Pe = (row, col, val * early)
    Y = Y0.copy()
    uY = np.zeros_like(Y)
    gains = np.ones_like(Y)
    mom = 0.5
    for it in range(iters):
        g, _ = orc.bh_gradient(Pe, Y, theta)
        L.orc_update(Y.reshape(-1), uY.reshape(-1), gains.reshape(-1), np.ascontiguousarray(g).reshape(-1), N, D, mom,
                     lr, msn, 0)
        L.orc_zero_mean(Y.reshape(-1), N, D)
        if it == stop:
            Pe = (row, col, Pe[2] / early)
        if it == mom_switch:
            mom = 0.8
    d = np.abs(Yg - Y).max()
    _note(f"  max |dY| {d:.2e} (extent {np.abs(Y).max():.3f})")
    assert d <= 1e-12 * np.abs(Y).max()


def test_program_barnes_hut_end_to_end(tmp_path):
    rng = np.random.default_rng(7)
    N = 1500
    X = rng.normal(0, 3, size=(6, 20))[rng.integers(0, 6, N)] + rng.normal(size=(N, 20))
    Y0 = rng.normal(0, 1e-4, size=(N, 2))
    _note(f"program nbody_algo=Barnes-Hut N={N} 300 iterations")
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        Y, loss = ft.fast_tsne(X, theta=0.5, nbody_algo="Barnes-Hut", perplexity=30, initialization=Y0,
                               max_iter=300, knn_algo="vp-tree", return_loss=True, load_affinities="save")
    finally:
        os.chdir(cwd)
    assert np.isfinite(Y).all() and Y.shape == (N, 2)
    row = np.fromfile(tmp_path / "P_row.dat", dtype=np.uint32)
    col = np.fromfile(tmp_path / "P_col.dat", dtype=np.uint32)
    val = np.fromfile(tmp_path / "P_val.dat", dtype=np.float64)
    # the oracle's loop from the program's own P and start
    L = orc.oracle()
    early, lr = 12.0, 200.0
    Pe = (row, col, val * early)
    Yr = Y0.copy()
    uY = np.zeros_like(Yr)
    gains = np.ones_like(Yr)
    mom = 0.5
    kl = {}
    for it in range(300):
        g, _ = orc.bh_gradient(Pe, Yr, 0.5)
        L.orc_update(Yr.reshape(-1), uY.reshape(-1), gains.reshape(-1), np.ascontiguousarray(g).reshape(-1), N, 2,
                     mom, lr, 5.0, 0)
        L.orc_zero_mean(Yr.reshape(-1), N, 2)
        if it == 250:
            Pe = (row, col, val)
            mom = 0.8
        if (it + 1) % 50 == 0:
            C = orc.bh_error(Pe, Yr, 0.5)
            kl[it] = C / early - math.log(early) if it < 250 else C
    got = loss[np.arange(49, 300, 50)]
    want = np.array([kl[i] for i in range(49, 300, 50)])
    _note(f"  KL {got} vs oracle {want}")
    assert abs(got[-1] - want[-1]) <= 5e-3 * want[-1]
