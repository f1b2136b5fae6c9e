"""The oracle against the reference's own outputs committed in tests/golden/ref_vectors.npz (written by
tests/golden/make_ref_golden.py from oracle/_ref, the reference compiled from its sources).

These checks need no reference sources: they keep the bit-exact pins of tests/test_oracle_ref.py,
tests/test_oracle_exact.py, tests/test_oracle_bh.py and tests/test_oracle_run.py alive in a tree where
oracle/_ref cannot be built.  Every comparison is bit for bit.
"""
import os

import numpy as np
import pytest

import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)


def csr(g, prefix):
    return (np.ascontiguousarray(g[prefix + "_row"], dtype=np.uint32),
            np.ascontiguousarray(g[prefix + "_col"], dtype=np.uint32),
            np.ascontiguousarray(g[prefix + "_val"], dtype=np.float64))


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_interpolate(g, p):
    assert np.array_equal(orc.interpolate(p, g["interp_y"], g[f"interp_sp{p}"]), g[f"interp_out{p}"])


def test_kernels(g):
    O = orc.oracle()
    for kind in (0, 1, 2):
        for j, df in enumerate((1.0, 0.5, 3.0)):
            k1 = np.array([O.orc_kernel_1d(kind, a, b, df) for a, b, _, _ in g["kern_args"]])
            k2 = np.array([O.orc_kernel_2d(kind, a, b, c, d, df) for a, b, c, d in g["kern_args"]])
            assert np.array_equal(k1, g[f"kern1_{kind}_{j}"])
            assert np.array_equal(k2, g[f"kern2_{kind}_{j}"])


def test_distances2similarities(g):
    O = orc.oracle()
    for j in range(3):
        perp, K, sq = g[f"d2s_cfg{j}"]
        D = g[f"d2s_D{j}"]
        for t in range(D.shape[0]):
            P = np.empty(int(K))
            b = O.orc_distances2similarities(np.ascontiguousarray(D[t]), P, int(K), -1, perp, -1.0, int(sq))
            assert b == g[f"d2s_beta{j}"][t]
            assert np.array_equal(P, g[f"d2s_P{j}"][t])
    P = np.empty(40)
    b = O.orc_distances2similarities_list(g["d2s_list_D"], P, 40, -1, 0.0, -1.0, 0, 3, g["d2s_list_pl"])
    assert b == g["d2s_list_beta"] and np.array_equal(P, g["d2s_list_P"])
    b = O.orc_distances2similarities(g["d2s_list_D"], P, 40, -1, -1.0, 2.0, 0)
    assert b == g["d2s_sigma_beta"] and np.array_equal(P, g["d2s_sigma_P"])


def test_symmetrize(g):
    out = orc.symmetrize(*csr(g, "sym_in"))
    for a, b in zip(out, (g["sym_row"], g["sym_col"], g["sym_val"])):
        assert np.array_equal(a, b)


def test_evaluate_error_fft(g):
    P = csr(g, "kl")
    got = [orc.kl(P, g["kl_Y2"], 1234.5, df=1.0), orc.kl(P, g["kl_Y2"], 1234.5, df=0.7), orc.kl(P, g["kl_Y1"], 99.0)]
    assert np.array_equal(np.array(got), g["kl_out"])


def test_exact_mode(g):
    assert np.array_equal(orc.squared_distances(g["ex_X"]), g["ex_dist"])
    Pcal = orc.gaussian_perplexity_dense(g["ex_X"], 8.0)
    assert np.array_equal(Pcal, g["ex_Pcal"])
    Pd, _ = orc.exact_symmetrize_normalize(Pcal)
    assert np.array_equal(Pd, g["ex_P"])
    assert np.array_equal(orc.exact_gradient(Pd, g["ex_Y"], df=1.0)[0], g["ex_grad1"])
    assert np.array_equal(orc.exact_gradient(Pd, g["ex_Y"], df=0.6)[0], g["ex_grad2"])
    kl = np.array([orc.exact_kl(Pd, g["ex_Y"], df=1.0), orc.exact_kl(Pd, g["ex_Y"], df=0.6)])
    assert np.array_equal(kl, g["ex_kl"])


def test_barnes_hut(g):
    P = csr(g, "bh")
    assert np.array_equal(orc.bh_gradient(P, g["bh_Y"], 0.5)[0], g["bh_grad"])
    assert orc.bh_error(P, g["bh_Y"], 0.5) == g["bh_err"]


def test_zero_mean(g):
    z = np.ascontiguousarray(g["zm_in"]).copy().reshape(-1)
    orc.oracle().orc_zero_mean(z, 200, 3)
    assert np.array_equal(z.reshape(200, 3), g["zm_out"])


def test_vptree_rows(g):
    idx, dist = orc.knn_exact(g["vp_X"], 15)
    row, col, val = orc.rows_from_knn(idx, dist, 5.0)
    assert np.array_equal(row, g["vp_row"]) and np.array_equal(col, g["vp_col"]) and np.array_equal(val, g["vp_val"])


RUN = dict(stop_lying_iter=5, mom_switch_iter=8, momentum=0.5, final_momentum=0.8, learning_rate=150.0,
           early_exag_coeff=12.0, no_momentum_during_exag=0, start_late_exag_iter=11, late_exag_coeff=3.0, df=1.0,
           max_step_norm=5.0, max_iter=14)


def test_run_loop(g):
    P = csr(g, "run")
    Y, c = orc.run_bh(P, g["run_Y0"], theta=0.5, **RUN)
    assert np.array_equal(Y, g["run_bh_Y"]) and np.array_equal(c, g["run_bh_costs"])
    Y, c = orc.run_bh(P, g["run_Y0"], theta=0.5, **dict(RUN, no_momentum_during_exag=1, learning_rate=40.0, max_iter=9))
    assert np.array_equal(Y, g["run_bhn_Y"]) and np.array_equal(c, g["run_bhn_costs"])
    Y, c = orc.run_exact(g["run_ex_P"], g["run_ex_Y0"], **dict(RUN, df=0.5, max_iter=12))
    assert np.array_equal(Y, g["run_ex_Y"]) and np.array_equal(c, g["run_ex_costs"])
