"""The drop-in program's file protocol and option handling, on the CPU.

The parameter/data file layout is the one TSNE::load_data reads (tsne.cpp:1915-1985); the result file
is TSNE::save_data's (tsne.cpp:2025-2036); the option rules are the reference wrapper's
(fast_tsne.py:177-256).  The reader below walks the file field by field, independently of the writer.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import fast_tsne as ft

INT_FIELDS_1 = ["no_dims", "max_iter", "stop_lying_iter", "mom_switch_iter"]
DBL_FIELDS_1 = ["momentum", "final_momentum", "learning_rate", "max_step_norm"]


def read_data_file(path):
    """Field-by-field reader of the data file (native endian, packed)."""
    buf = open(path, "rb").read()
    off = 0
    out = {}

    def take(fmt, name):
        nonlocal off
        (v,) = struct.unpack_from("=" + fmt, buf, off)
        off += struct.calcsize("=" + fmt)
        out[name] = v

    take("i", "n"), take("i", "d"), take("d", "theta"), take("d", "perplexity")
    if out["perplexity"] == 0:
        take("i", "L")
        out["perplexity_list"] = list(np.frombuffer(buf, dtype=np.float64, count=out["L"], offset=off))
        off += 8 * out["L"]
    for f in INT_FIELDS_1:
        take("i", f)
    for f in DBL_FIELDS_1:
        take("d", f)
    for fmt, f in [("i", "K"), ("d", "sigma"), ("i", "nbody_algo"), ("i", "knn_algo"), ("d", "early_exag_coeff"),
                   ("i", "no_momentum_during_exag"), ("i", "n_trees"), ("i", "search_k"),
                   ("i", "start_late_exag_iter"), ("d", "late_exag_coeff"), ("i", "nterms"),
                   ("d", "intervals_per_integer"), ("i", "min_num_intervals")]:
        take(fmt, f)
    nd = out["n"] * out["d"]
    out["X"] = np.frombuffer(buf, dtype=np.float64, count=nd, offset=off).reshape(out["n"], out["d"])
    off += 8 * nd
    take("i", "seed"), take("d", "df"), take("i", "load_affinities")
    rest = (len(buf) - off) // 8
    out["Y"] = np.frombuffer(buf, dtype=np.float64, count=rest, offset=off) if rest else None
    return out


def _write(path, X, Y0, perplexity=30.0, plist=None, **over):
    kw = dict(theta=0.5, perplexity=perplexity, perplexity_list=plist, map_dims=2, max_iter=7,
              stop_early_exag_iter=3, mom_switch_iter=4, momentum=0.25, final_momentum=0.75, learning_rate=321.5,
              max_step_norm=4.5, K=-1, sigma=-1.0, nbody_algo=2, knn_algo=2, early_exag_coeff=11.0,
              no_momentum_during_exag=0, n_trees=17, search_k=999, start_late_exag_iter=-1, late_exag_coeff=-1.0,
              nterms=3, intervals_per_integer=1.0, min_num_intervals=50, seed=123, df=0.75, load_affinities=0,
              initialization=Y0)
    kw.update(over)
    ft.write_data_file(path, X, **kw)
    return kw


def test_data_file_fields_in_load_data_order(tmp_path):
    rng = np.random.default_rng(1)
    X = rng.normal(size=(37, 5))
    Y0 = rng.normal(size=(37, 2))
    kw = _write(tmp_path / "d.dat", X, Y0)
    r = read_data_file(tmp_path / "d.dat")
    assert (r["n"], r["d"]) == (37, 5)
    for k in ["theta", "perplexity", "max_iter", "momentum", "final_momentum", "learning_rate", "max_step_norm",
              "K", "sigma", "nbody_algo", "knn_algo", "early_exag_coeff", "n_trees", "search_k", "late_exag_coeff",
              "nterms", "intervals_per_integer", "min_num_intervals", "seed", "df", "load_affinities"]:
        assert r[k] == kw[k], k
    assert r["no_dims"] == 2 and r["stop_lying_iter"] == 3 and r["mom_switch_iter"] == 4
    np.testing.assert_array_equal(r["X"], X)
    np.testing.assert_array_equal(r["Y"], Y0.reshape(-1))


def test_data_file_with_perplexity_list_and_random_init(tmp_path):
    X = np.arange(24, dtype=np.float64).reshape(8, 3)
    _write(tmp_path / "d.dat", X, "random", perplexity=0.0, plist=[5.0, 10.0, 2.5])
    r = read_data_file(tmp_path / "d.dat")
    assert r["perplexity"] == 0 and r["L"] == 3 and r["perplexity_list"] == [5.0, 10.0, 2.5]
    assert r["Y"] is None
    np.testing.assert_array_equal(r["X"], X)


def test_result_file_reader_masks_uncomputed_losses(tmp_path):
    n, md, mi = 5, 2, 120
    Y = np.arange(n * md, dtype=np.float64)
    costs = np.linspace(1, 2, mi)
    with open(tmp_path / "r.dat", "wb") as f:
        f.write(struct.pack("=ii", n, md) + Y.tobytes() + struct.pack("=i", mi) + costs.tobytes())
    Yr, loss = ft.read_result_file(tmp_path / "r.dat", mi)
    np.testing.assert_array_equal(Yr, Y.reshape(n, md))
    keep = (np.arange(1, mi + 1) % 50) == 0
    np.testing.assert_array_equal(loss[keep], costs[keep])
    assert np.isnan(loss[~keep]).all()


def test_option_resolution_follows_the_reference_wrapper():
    X = np.random.default_rng(0).normal(size=(6000, 4))
    base = dict(theta=0.5, perplexity=30, map_dims=2, early_exag_coeff=12, learning_rate="auto",
                start_late_exag_iter="auto", stop_early_exag_iter=250, late_exag_coeff=-1, initialization="random",
                seed=-1, perplexity_list=None, sigma=-1, K=-1, search_k=None, n_trees=50, nbody_algo="FFT",
                knn_algo="annoy", load_affinities=None, nthreads=-1, max_step_norm=5, no_momentum_during_exag=False)
    r = ft._resolve(X, **base)
    assert r["learning_rate"] == 500.0 and r["start_late_exag_iter"] == -1 and r["search_k"] == 3 * 30 * 50
    assert r["nbody_algo"] == 2 and r["knn_algo"] == 1 and r["load_affinities"] == 0 and r["nthreads"] == 0
    r = ft._resolve(X[:100], **{**base, "late_exag_coeff": 4, "max_step_norm": "none", "knn_algo": "vp-tree",
                                "load_affinities": "save"})
    assert r["learning_rate"] == 200.0 and r["start_late_exag_iter"] == 250 and r["max_step_norm"] == -1
    assert r["knn_algo"] == 2 and r["load_affinities"] == 2
    r = ft._resolve(X, **{**base, "perplexity_list": [10, 40]})
    assert r["perplexity"] == 0 and r["search_k"] == 3 * 40 * 50
    r = ft._resolve(X, **{**base, "sigma": 2.0, "K": 15})
    assert r["perplexity"] == -1 and r["search_k"] == 15 * 50
    with pytest.raises(ValueError):
        ft._resolve(X, **{**base, "nbody_algo": "exact"})


def _run_binary(args, cwd):
    return subprocess.run([ft.BINARY] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_binary_rejects_a_wrong_version(tmp_path):
    r = _run_binary(["1.2.0"], tmp_path)
    assert r.returncode != 0 and "wrong version 1.2.0" in r.stdout


def _echoed_settings(out):
    line = next(ln for ln in out.splitlines() if ln.startswith("[fast_tsne] settings:"))
    return {k: v for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}


def test_binary_reads_the_data_file(tmp_path):
    """The program logs what it parsed before it needs a GPU; those values are the file's."""
    rng = np.random.default_rng(2)
    X = rng.normal(size=(60, 4))
    kw = _write(tmp_path / "d.dat", X, rng.normal(size=(60, 2)) * 1e-4, perplexity=5.0, max_iter=3)
    r = _run_binary([ft.VERSION, "d.dat", "r.dat", "1"], tmp_path)
    out = r.stdout
    assert "[fast_tsne] data: N=60 D=4" in out
    got = _echoed_settings(out)
    for k in ["theta", "perplexity", "learning_rate", "max_step_norm", "early_exag_coeff", "df", "sigma",
              "intervals_per_integer"]:
        assert float(got[k]) == kw[k], k
    for k in ["max_iter", "K", "nbody_algo", "knn_algo", "n_trees", "search_k", "nterms", "min_num_intervals"]:
        assert int(got[k]) == kw[k], k
    assert int(got["no_dims"]) == 2 and int(got["stop_lying_iter"]) == 3 and int(got["seed"]) == 123
    assert "[fast_tsne] initial embedding: given" in out
    if r.returncode != 0:  # no GPU here: the program stops with a message instead of computing on the CPU
        assert "no GPU device is visible" in r.stderr


def test_binary_rejects_invalid_settings(tmp_path):
    X = np.random.default_rng(3).normal(size=(40, 3))
    _write(tmp_path / "d.dat", X, "random", perplexity=5.0, nbody_algo=3)
    r = _run_binary([ft.VERSION, "d.dat", "r.dat"], tmp_path)
    assert r.returncode != 0 and "undefined algorithm" in r.stderr
    # the FFT setting limits no_dims to 1 or 2 whatever theta is (tsne.cpp:1954-1957)
    _write(tmp_path / "d.dat", X, "random", perplexity=5.0, theta=0.0, map_dims=3)
    r = _run_binary([ft.VERSION, "d.dat", "r.dat"], tmp_path)
    assert r.returncode != 0 and "needs no_dims 1 or 2" in r.stderr


def test_binary_accepts_the_exact_and_barnes_hut_modes(tmp_path):
    """theta = 0 runs the exact mode, nbody_algo = 1 with theta > 0 the Barnes-Hut mode (both also in 3
    output dimensions); without a GPU the program stops before computing."""
    X = np.random.default_rng(4).normal(size=(40, 3))
    for dims, algo, theta in [(2, 2, 0.0), (3, 1, 0.0), (3, 1, 0.5), (1, 1, 0.5)]:
        _write(tmp_path / "d.dat", X, "random", perplexity=5.0, theta=theta, map_dims=dims, nbody_algo=algo,
               max_iter=3)
        r = _run_binary([ft.VERSION, "d.dat", "r.dat"], tmp_path)
        if r.returncode != 0:
            assert "no GPU device is visible" in r.stderr, r.stderr
        else:
            Y, _ = ft.read_result_file(tmp_path / "r.dat", 3)
            assert Y.shape == (40, dims) and np.isfinite(Y).all()


# ---------------------------------------------------------------- against the reference's own parser and writer

import orc  # noqa: E402

needs_ref = pytest.mark.skipif(not orc.have_ref(), reason="oracle/_ref not built (reference sources absent)")

REF_INTS = ["n", "d", "no_dims", "max_iter", "stop_lying_iter", "mom_switch_iter", "K", "nbody_algo", "knn_algo",
            "no_momentum_during_exag", "n_trees", "search_k", "start_late_exag_iter", "nterms", "min_num_intervals",
            "seed", "load_affinities", "skip_random_init", "perplexity_list_length"]
REF_DBLS = ["theta", "perplexity", "momentum", "final_momentum", "learning_rate", "max_step_norm", "sigma",
            "early_exag_coeff", "late_exag_coeff", "intervals_per_integer", "df"]


def ref_load(path, n, d, nd, L):
    ints = np.zeros(19, dtype=np.int32)
    dbls = np.zeros(11)
    X = np.zeros(n * d)
    Y = np.full(n * nd, np.nan)
    pl = np.zeros(max(L, 1))
    rc = orc.ref().ref_load_data(str(path).encode(), ints, dbls, X.ctypes.data, Y.ctypes.data, pl.ctypes.data)
    assert rc == 0
    return dict(zip(REF_INTS, ints.tolist())), dict(zip(REF_DBLS, dbls.tolist())), X.reshape(n, d), Y, pl[:L]


VARIANTS = {
    "perplexity_given_init": dict(perplexity=30.0, plist=None, init="given"),
    "perplexity_list_random_init": dict(perplexity=0.0, plist=[5.0, 10.0, 2.5], init="random"),
    "sigma_K": dict(perplexity=-1.0, plist=None, init="random", sigma=2.0, K=15),
    "load_affinities_late_exag": dict(perplexity=12.0, plist=None, init="given", load_affinities=1, df=0.6,
                                      start_late_exag_iter=5, late_exag_coeff=4.0, no_momentum_during_exag=1,
                                      nbody_algo=1, theta=0.3, map_dims=1),
}


@needs_ref
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_reference_load_data_reads_the_wrapper_file(tmp_path, variant):
    """The data file written by fast_tsne.write_data_file, parsed by the reference's own TSNE::load_data
    (tsne.cpp:1895-2022, compiled from its sources): every field, X, the initial Y and the perplexity list."""
    v = dict(VARIANTS[variant])
    rng = np.random.default_rng(hash(variant) % 1000)
    n, d = 41, 6
    nd = v.pop("map_dims", 2)
    X = rng.normal(size=(n, d))
    init = rng.normal(size=(n, nd)) if v.pop("init") == "given" else "random"
    perplexity, plist = v.pop("perplexity"), v.pop("plist")
    kw = _write(tmp_path / "d.dat", X, init, perplexity=perplexity, plist=plist, map_dims=nd, **v)
    L = len(plist) if plist else 0
    ints, dbls, Xr, Yr, plr = ref_load(tmp_path / "d.dat", n, d, nd, L)
    assert (ints["n"], ints["d"], ints["no_dims"]) == (n, d, nd)
    assert ints["stop_lying_iter"] == kw["stop_early_exag_iter"]
    for k in ["max_iter", "mom_switch_iter", "K", "nbody_algo", "knn_algo", "no_momentum_during_exag", "n_trees",
              "search_k", "start_late_exag_iter", "nterms", "min_num_intervals", "seed", "load_affinities"]:
        assert ints[k] == kw[k], k
    for k in REF_DBLS:
        assert dbls[k] == kw[k], k
    assert ints["skip_random_init"] == (0 if isinstance(init, str) else 1)
    np.testing.assert_array_equal(Xr, X)
    if not isinstance(init, str):
        np.testing.assert_array_equal(Yr, init.reshape(-1))
    assert ints["perplexity_list_length"] == L and list(plr) == (plist or [])
    # the drop-in program parses the same file to the same values
    r = _run_binary([ft.VERSION, "d.dat", "r.dat"], tmp_path)
    got = _echoed_settings(r.stdout)
    for k in ["max_iter", "K", "nbody_algo", "knn_algo", "n_trees", "search_k", "nterms", "min_num_intervals",
              "start_late_exag_iter", "load_affinities", "no_momentum_during_exag", "no_dims"]:
        assert int(got[k]) == ints[k], k
    for k in ["theta", "perplexity", "learning_rate", "max_step_norm", "early_exag_coeff", "late_exag_coeff",
              "df", "sigma", "intervals_per_integer", "momentum", "final_momentum"]:
        assert float(got[k]) == pytest.approx(dbls[k], rel=1e-5), k


@needs_ref
@pytest.mark.parametrize("max_iter", [120, 100, 1])
def test_result_file_from_the_reference_save_data(tmp_path, max_iter):
    """A result file written by the reference's TSNE::save_data (tsne.cpp:2024-2038) read by
    fast_tsne.read_result_file: the embedding exactly, the losses at every 50th and the last iteration,
    NaN elsewhere (the wrapper's rule, fast_tsne.py:328)."""
    n, d = 13, 2
    rng = np.random.default_rng(max_iter)
    Y = rng.normal(size=n * d)
    costs = rng.uniform(1, 3, size=max_iter)
    orc.ref().ref_save_data(str(tmp_path / "r.dat").encode(), Y.copy(), costs.copy(), n, d, max_iter)
    Yr, loss = ft.read_result_file(tmp_path / "r.dat", max_iter)
    np.testing.assert_array_equal(Yr, Y.reshape(n, d))
    keep = (np.arange(1, max_iter + 1) % 50) == 0
    np.testing.assert_array_equal(loss[keep], costs[keep])
    assert np.isnan(loss[~keep]).all()
