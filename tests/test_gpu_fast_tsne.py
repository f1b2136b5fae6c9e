"""The drop-in program and its affinity build on the GPU, against the oracle.

* fitsne_knn_exact vs the oracle's exact kNN (the reference's VP-tree search returns the exact
  neighbours; its distance is vptree.h:97-107): same indices, same distances.
* bin/fast_tsne with load_affinities='save' writes P_row/P_col/P_val.dat; they must equal the oracle's
  computeGaussianPerplexity + symmetrizeMatrix + normalisation (tsne.cpp:151-330) on the same data.
* The affinity modes: a perplexity list (K = 3 max, rows averaged) and a fixed sigma with K neighbours on
  un-normalised data, against the oracle (same rows and columns, values 1e-12); load_affinities='load'
  reproduces the saving run bit for bit.
* The optimisation schedule (early and late exaggeration, momentum switch, no_momentum_during_exag, 1D
  with df != 1): the program (FP64 plain path) against the oracle's run of the same loop, 1e-9.
* C1 end to end (N = 10,000, 1000 iterations): the program's final KL within 0.5% of the oracle's run
  of TSNE::run on the same data and initialisation (tests/golden/c1_oracle_kl.json, made by
  tests/golden/make_c1_golden.py), in the FP32 form and the FP64 plain path.
"""
import json
import os
import sys

import numpy as np
import pytest

import fast_tsne as ft
import fitsne_gpu as fg
import orc
import synth

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_c1_golden as c1  # noqa: E402

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[fast_tsne test] {msg}", flush=True)


def test_knn_exact_matches_the_oracle():
    rng = np.random.default_rng(11)
    X = np.concatenate([rng.normal(c, 1.0, size=(700, 20)) for c in rng.normal(0, 3, size=(4, 20))])
    X[5] = X[9]  # one exact duplicate: a zero distance, ties broken by the smaller index
    K = 45
    _note(f"knn N={X.shape[0]} D=20 K={K}")
    idx, dist = fg.knn_exact(X, K)
    ridx, rdist = orc.knn_exact(X, K)
    np.testing.assert_array_equal(idx, ridx)
    np.testing.assert_array_equal(dist, rdist)
    assert idx[5, 0] == 9 and dist[5, 0] == 0.0 and idx[9, 0] == 5


def _run_program(tmp_path, X, Y0, max_iter, precision=None, **kw):
    env_before = os.environ.get("FITSNE_PRECISION")
    cwd = os.getcwd()
    try:
        if precision:
            os.environ["FITSNE_PRECISION"] = precision
        os.chdir(tmp_path)
        return ft.fast_tsne(X, perplexity=30, max_iter=max_iter, knn_algo="vp-tree", initialization=Y0,
                            return_loss=True, **kw)
    finally:
        os.chdir(cwd)
        if env_before is None:
            os.environ.pop("FITSNE_PRECISION", None)
        else:
            os.environ["FITSNE_PRECISION"] = env_before


def test_program_affinities_match_the_oracle(tmp_path):
    X, _ = synth.c1_data(N=2000, D=30, k=6, seed=5)
    Y0 = c1.c1_init(X)
    _note("program N=2000 load_affinities=save max_iter=2")
    _run_program(tmp_path, X, Y0, max_iter=2, load_affinities="save")
    row = np.fromfile(tmp_path / "P_row.dat", dtype=np.uint32)
    col = np.fromfile(tmp_path / "P_col.dat", dtype=np.uint32)
    val = np.fromfile(tmp_path / "P_val.dat", dtype=np.float64)
    rrow, rcol, rval = c1.c1_affinities(X)
    np.testing.assert_array_equal(row, rrow)
    np.testing.assert_array_equal(col, rcol)
    np.testing.assert_allclose(val, rval, rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_c1_final_kl_within_half_percent_of_the_oracle(tmp_path, precision):
    gold = json.load(open(os.path.join(HERE, "golden", "c1_oracle_kl.json")))
    X, _ = synth.c1_data()
    Y0 = c1.c1_init(X)
    _note(f"program C1 N=10000 1000 iterations {precision}")
    Y, loss = _run_program(tmp_path, X, Y0, max_iter=1000, precision=precision)
    kl = loss[np.arange(49, 1000, 50)]
    _note(f"{precision}: KL {kl[[0, 4, 9, -1]]} vs oracle {np.array(gold['kl'])[[0, 4, 9, -1]]}")
    assert np.isfinite(Y).all() and Y.shape == (10000, 2)
    assert abs(kl[-1] - gold["kl"][-1]) / gold["kl"][-1] < 5e-3


@pytest.mark.parametrize("dims,df,nomom", [(1, 0.6, 0), (2, 1.0, 1)])
def test_program_schedule_tracks_the_oracle_run(tmp_path, dims, df, nomom):
    """The program's optimisation schedule (tsne.cpp:389-577) against the oracle's run of the same loop
    (orc_run_fft) on the same P and initial Y, with the FP64 plain path: early exaggeration until
    iteration 30, the momentum switch at 30, late exaggeration x3 from 60, in 1D with df = 0.6 and in 2D
    with no_momentum_during_exag.  Any switch a step early or late moves the KL entries far more than
    the bound."""
    N = 1500
    X, _ = synth.c1_data(N=N, D=30, k=5, seed=9)
    Y0 = np.random.default_rng(3).normal(0, 1e-4, size=(N, dims))
    sched = dict(stop_early_exag_iter=30, mom_switch_iter=30, start_late_exag_iter=60, late_exag_coeff=3.0,
                 early_exag_coeff=12.0, learning_rate=200.0, max_step_norm=5.0)
    _note(f"program schedule N={N} dims={dims} df={df} no_momentum_during_exag={nomom}")
    Y, loss = _run_program(tmp_path, X, Y0, max_iter=120, precision="fp64", map_dims=dims, df=df,
                           no_momentum_during_exag=bool(nomom), **sched)
    P = c1.c1_affinities(X)
    sched["stop_lying_iter"] = sched.pop("stop_early_exag_iter")
    Yr, costs = orc.run_fft(P, Y0, max_iter=120, df=df, no_momentum_during_exag=nomom, **sched)
    # the result reader keeps the multiples of 50 only (uncomputed entries NaN, as the reference's)
    at = np.nonzero(np.isfinite(loss))[0]
    np.testing.assert_array_equal(at, [49, 99])
    assert (costs[at] > 0).all()
    r_kl = np.abs(loss[at] - costs[at]) / costs[at]
    r_y = np.linalg.norm(Y.reshape(Yr.shape) - Yr) / np.linalg.norm(Yr)
    _note(f"KL entries at {list(at)}: max rel diff {r_kl.max():.2e}; final Y rel diff {r_y:.2e}")
    # measured 1e-13 (KL) and 6e-12 (Y): the plain path's gradient agrees with the oracle's to ~1e-12
    assert r_kl.max() <= 1e-9
    assert r_y <= 1e-9


def _oracle_P_modes(X, K, perplexity, sigma=-1.0, plist=None):
    """The oracle's computeGaussianPerplexity (list and fixed-sigma modes, tsne.cpp:1643-1726 with the
    list overload :1474-1500) + symmetrizeMatrix + normalisation on the program's prepared data: X
    zero-mean, and scaled by its largest magnitude only when a perplexity is used (tsne.cpp:153-164)."""
    N, D = X.shape
    Xc = np.ascontiguousarray(X, dtype=np.float64).copy()
    orc.oracle().orc_zero_mean(Xc.reshape(-1), N, D)
    if perplexity > 0 or plist is not None:
        Xc /= np.abs(Xc).max()
    idx, dist = orc.knn_exact(Xc, K, nthreads=os.cpu_count() or 1)
    row, col, val = orc.rows_from_knn(idx, dist, perplexity, sigma=sigma, plist=plist)
    return orc.normalize(*orc.symmetrize(row, col, val))


@pytest.mark.parametrize("mode", ["perplexity_list", "sigma"])
def test_program_affinity_modes_match_the_oracle(tmp_path, mode):
    X, _ = synth.c1_data(N=1800, D=30, k=6, seed=7)
    Y0 = c1.c1_init(X)
    if mode == "perplexity_list":
        kw = dict(perplexity_list=[10, 25])
        want = _oracle_P_modes(X, 75, 0.0, plist=np.array([10.0, 25.0]))
    else:
        # the data stay un-normalised in this mode: within-cluster squared distances are ~60
        kw = dict(sigma=5.0, K=40)
        want = _oracle_P_modes(X, 40, -1.0, sigma=5.0)
    _note(f"program N=1800 affinities, {mode}")
    _run_program(tmp_path, X, Y0, max_iter=2, load_affinities="save", **kw)
    row = np.fromfile(tmp_path / "P_row.dat", dtype=np.uint32)
    col = np.fromfile(tmp_path / "P_col.dat", dtype=np.uint32)
    val = np.fromfile(tmp_path / "P_val.dat", dtype=np.float64)
    np.testing.assert_array_equal(row, want[0])
    np.testing.assert_array_equal(col, want[1])
    np.testing.assert_allclose(val, want[2], rtol=1e-12, atol=0)


def test_program_loaded_affinities_reproduce_the_run(tmp_path):
    """load_affinities='load' reads P_row/P_col/P_val.dat (tsne.cpp:232-265) instead of building P: a run
    on those files is the run that saved them, bit for bit (the FP32 step is deterministic)."""
    X, _ = synth.c1_data(N=1500, D=30, k=5, seed=11)
    Y0 = c1.c1_init(X)
    _note("program N=1500 save, then load")
    Y1 = _run_program(tmp_path, X, Y0, max_iter=40, load_affinities="save")[0]
    Y2 = _run_program(tmp_path, X, Y0, max_iter=40, load_affinities="load")[0]
    np.testing.assert_array_equal(Y1, Y2)
