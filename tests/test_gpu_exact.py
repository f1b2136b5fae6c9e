"""Exact t-SNE (theta = 0) on the GPU against the oracle.

The oracle's exact functions are bit-exact restatements of the reference's (tests/test_oracle_exact.py
pins them to oracle/_ref and to dKL/dY = 4 dC).  Here, on the same inputs:
* fitsne_exact_affinities vs the dense computeGaussianPerplexity + symmetrisation + normalisation
  (perplexity, perplexity-list and fixed-sigma modes): rtol 1e-11 (the device exp and the tree sums
  differ from the host's in the last bits), exactly symmetric;
* the exact gradient (computeExactGradient) in 1-4 dimensions, df = 1 and df != 1: ||d||/||g|| <= 1e-12,
  Z within 1e-13; bitwise repeatable;
* evaluateError: within 1e-12 relative;
* 25 optimisation steps (gains, momentum, clip, zeroMean, exaggeration switch) vs the oracle's loop:
  max |dY| <= 1e-9 of the embedding's extent;
* the drop-in program with theta = 0 end to end (N = 1000, 300 iterations) vs the oracle's loop: final KL
  within 0.5%; P.dat written by load_affinities='save' equals the oracle's P; a run from P.dat
  (load_affinities='load') gives the same embedding bit for bit;
* the FFT gradient's interpolation error, measured against the exact gradient with the same P at
  N = 20,000 (the exact path as a mid-N oracle): FP64 plain and FP32 forms.
"""
import math
import os

import numpy as np
import pytest

import fast_tsne as ft
import fitsne_gpu as fg
import orc
import synth

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[exact test] {msg}", flush=True)


def _data(N, D, seed, k=5):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 4, size=(k, D))
    X = c[rng.integers(0, k, N)] + rng.normal(size=(N, D))
    X -= X.mean(0)
    return X / np.abs(X).max()


def _ctx(D, df=1.0):
    return fg.Context(no_dims=D, precision=fg.FP64_EXACT, df=df)


@pytest.mark.parametrize("mode", ["perplexity", "list", "sigma"])
def test_exact_affinities_match_the_oracle(mode):
    X = _data(1200, 15, 1)
    kw = {"perplexity": dict(perplexity=30.0), "list": dict(perplexity=0.0, perplexity_list=[10.0, 50.0]),
          "sigma": dict(perplexity=-1.0, sigma=0.2)}[mode]
    _note(f"affinities N=1200 D=15 {mode}")
    c = _ctx(2)
    P, s = c.exact_affinities(X, **kw)
    c.close()
    ref_rows = orc.gaussian_perplexity_dense(X, kw["perplexity"], kw.get("sigma", -1.0), kw.get("perplexity_list"))
    Pr, sr = orc.exact_symmetrize_normalize(ref_rows)
    np.testing.assert_array_equal(P, P.T)
    assert abs(s - sr) <= 1e-12 * sr
    np.testing.assert_allclose(P, Pr, rtol=1e-11, atol=1e-300)


@pytest.mark.parametrize("N,D,df", [(2500, 2, 1.0), (2000, 1, 0.5), (1500, 3, 1.0), (1000, 4, 1.3), (777, 2, 0.8)])
def test_exact_gradient_matches_the_oracle(N, D, df):
    X = _data(N, 12, 2)
    P, _ = orc.exact_P(X, 25.0)
    Y = np.random.default_rng(3).normal(0, 5, size=(N, D))
    _note(f"gradient N={N} D={D} df={df}")
    c = _ctx(D, df)
    c.set_P_dense(P)
    c.set_Y(Y)
    g, z = c.gradient()
    g2, z2 = c.gradient()
    kl = c.kl()
    c.close()
    gr, zr = orc.exact_gradient(P, Y, df)
    np.testing.assert_array_equal(g, g2)
    assert z == z2
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    _note(f"  rel err {err:.2e}, Z {z:.12e} vs {zr:.12e}")
    assert err <= 1e-12
    assert abs(z - zr) <= 1e-13 * zr
    klr = orc.exact_kl(P, Y, df)
    assert abs(kl - klr) <= 1e-12 * abs(klr)


def _oracle_steps(P, Y, iters, early, stop, mom_switch, lr, msn, df, kl_every=50):
    L = orc.oracle()
    N, D = Y.shape
    P = P * early
    Y = Y.copy()
    uY = np.zeros_like(Y)
    gains = np.ones_like(Y)
    mom = 0.5
    costs = {}
    for it in range(iters):
        g, _ = orc.exact_gradient(P, Y, df)
        L.orc_update(Y.reshape(-1), uY.reshape(-1), gains.reshape(-1), np.ascontiguousarray(g).reshape(-1), N, D, mom,
                     lr, msn, 0)
        L.orc_zero_mean(Y.reshape(-1), N, D)
        if it == stop:
            P = P / early
        if it == mom_switch:
            mom = 0.8
        if (it + 1) % kl_every == 0 or it == iters - 1:
            C = orc.exact_kl(P, Y, df)
            if it < stop:
                C = C / early - math.log(early)
            costs[it] = C
    return Y, costs


def test_exact_steps_match_the_oracle():
    N, D, df = 900, 2, 1.0
    X = _data(N, 10, 4)
    P, _ = orc.exact_P(X, 20.0)
    Y0 = np.random.default_rng(5).normal(0, 1e-2, size=(N, D))
    early, stop, mom_switch, lr, msn, iters = 12.0, 10, 15, 200.0, 5.0, 25
    _note(f"steps N={N} iters={iters}")
    c = _ctx(D, df)
    c.set_P_dense(P)
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
    Yr, _ = _oracle_steps(P, Y0, iters, early, stop, mom_switch, lr, msn, df)
    ext = np.abs(Yr).max()
    d = np.abs(Yg - Yr).max()
    _note(f"  max |dY| {d:.2e} (extent {ext:.3f})")
    assert d <= 1e-9 * ext


def _run_program(tmp_path, X, Y0, **kw):
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        return ft.fast_tsne(X, theta=0.0, perplexity=30, initialization=Y0, return_loss=True, **kw)
    finally:
        os.chdir(cwd)


def test_program_exact_mode_end_to_end(tmp_path):
    N, iters = 1000, 300
    X, _ = synth.c1_data(N=N, D=20, k=6, seed=9)
    Y0 = np.random.default_rng(10).normal(0, 1e-4, size=(N, 2))
    _note(f"program theta=0 N={N} {iters} iterations, P.dat saved")
    Ys, loss = _run_program(tmp_path, X, Y0, max_iter=iters, load_affinities="save")
    # the oracle: the program's data preparation (tsne.cpp:153-165), P, then the loop
    Xn = X - X.mean(0)
    Xn /= np.abs(Xn).max()
    P, _ = orc.exact_P(Xn, 30.0)
    Pd = np.fromfile(tmp_path / "P.dat", dtype=np.float64).reshape(N, N)
    np.testing.assert_allclose(Pd, P, rtol=1e-11, atol=1e-300)
    Yr, costs = _oracle_steps(Pd, Y0, iters, 12.0, 250, 250, 200.0, 5.0, 1.0)
    kl = loss[np.arange(49, iters, 50)]
    klr = np.array([costs[i] for i in range(49, iters, 50)])
    _note(f"  KL {kl} vs oracle {klr}")
    assert abs(kl[-1] - klr[-1]) <= 5e-3 * klr[-1]
    _note("program theta=0 from P.dat")
    Yl, loss_l = _run_program(tmp_path, X, Y0, max_iter=iters, load_affinities="load")
    np.testing.assert_array_equal(Yl, Ys)


def test_fft_interpolation_error_against_the_exact_gradient():
    N = 20000
    P, Y, _ = synth.small_mixture_workload(N=N, D=2, k=10, per_point=30, span=40.0, seed=11)
    row, col, val = P
    dense = np.zeros((N, N))
    dense[np.repeat(np.arange(N), np.diff(row.astype(np.int64))), col.astype(np.int64)] = val
    _note(f"exact vs FFT gradient N={N} nnz={val.size}")
    c = _ctx(2)
    c.set_P_dense(dense)
    c.set_Y(Y)
    ge, ze = c.gradient()
    c.close()
    del dense
    out = {}
    for name, prec in [("fp64", fg.FP64_PLAIN), ("fp32", fg.FP32)]:
        c = fg.Context(no_dims=2, precision=prec)
        c.set_P(row, col, val)
        c.set_Y(Y)
        g, z = c.gradient()
        c.close()
        out[name] = (np.linalg.norm(g - ge) / np.linalg.norm(ge), abs(z - ze) / ze)
        _note(f"  {name}: gradient rel err {out[name][0]:.3e}, Z rel err {out[name][1]:.3e}")
    # interpolation error of p = 3 nodes per box (box width 0.8 here): 2.0e-2 on the gradient, 8.8e-4 on
    # Z; the oracle's FFT gradient has the same size of error against its exact gradient on this kind
    # of data (2.3e-2 and 8.8e-4 at N = 4000)
    assert out["fp64"][0] < 4e-2 and out["fp64"][1] < 2e-3
    # the FP32 form adds nothing visible on top of the interpolation error
    assert abs(out["fp32"][0] - out["fp64"][0]) <= 1e-3 * max(out["fp64"][0], 1e-6) + 1e-5
