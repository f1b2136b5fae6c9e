"""Writes tests/golden/ref_vectors.npz: inputs and the reference's own outputs for the oracle's pinned
functions, so that tests/test_golden.py keeps those pins where the reference sources are absent.

Every output comes from the reference's code compiled from its sources by oracle/Makefile
(oracle/_ref/libfitsne_ref.so for the functions, oracle/_ref/libfitsne_ref_run.so for TSNE::run's loop);
the inputs are small seeded arrays stored beside them.  Nothing of the reference's source is stored.

Usage (from the repository root, after `make -C oracle`): python tests/golden/make_ref_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))

import orc  # noqa: E402
from datagen import blobs, mixture, random_csr  # noqa: E402

OUT = os.path.join(HERE, "ref_vectors.npz")

RUN_BASE = dict(perplexity=5.0, theta=0.5, stop_lying_iter=5, mom_switch_iter=8, momentum=0.5, final_momentum=0.8,
                learning_rate=150.0, early_exag_coeff=12.0, no_momentum_during_exag=0, start_late_exag_iter=11,
                late_exag_coeff=3.0, df=1.0, max_step_norm=5.0, nbody_algorithm=1, max_iter=14)


def main():
    if not (orc.have_ref() and orc.have_ref_run()):
        raise SystemExit("oracle/_ref is not built: run make -C oracle with the reference sources present")
    R = orc.ref()
    v = {}

    # interpolate (nbodyfft.cpp:310-336)
    rng = np.random.default_rng(1)
    y = np.concatenate([rng.uniform(0, 1, 300), [0.0, 1.0, -0.01, 1.02]])
    v["interp_y"] = y
    for p in (2, 3, 4, 5):
        sp = np.empty(p)
        h = 1.0 / p
        sp[0] = h / 2
        for i in range(1, p):
            sp[i] = sp[i - 1] + h
        v[f"interp_sp{p}"] = sp
        v[f"interp_out{p}"] = orc.interpolate(p, y, sp, lib="ref")

    # kernels (tsne.cpp:69-94)
    args = rng.normal(scale=5, size=(60, 4))
    v["kern_args"] = args
    for kind in (0, 1, 2):
        for j, df in enumerate((1.0, 0.5, 3.0)):
            v[f"kern1_{kind}_{j}"] = np.array([R.ref_kernel_1d(kind, a, b, df) for a, b, _, _ in args])
            v[f"kern2_{kind}_{j}"] = np.array([R.ref_kernel_2d(kind, a, b, c, d, df) for a, b, c, d in args])

    # distances2similarities (tsne.cpp:1394-1500)
    for j, (perp, K, sq) in enumerate(((30.0, 90, 0), (5.0, 15, 1), (50.0, 150, 0))):
        D = np.sort(rng.gamma(2.0, 1.0 + j, size=(4, K)), axis=1)
        P = np.empty_like(D)
        beta = np.array([R.ref_distances2similarities(D[t].copy(), P[t], K, -1, perp, -1.0, sq) for t in range(4)])
        v[f"d2s_D{j}"], v[f"d2s_P{j}"], v[f"d2s_beta{j}"] = D, P, beta
        v[f"d2s_cfg{j}"] = np.array([perp, K, sq])
    D = np.sort(rng.gamma(2.0, 1.0, size=40))
    pl = np.array([5.0, 15.0, 10.0])
    P = np.empty(40)
    v["d2s_list_D"], v["d2s_list_pl"] = D, pl
    v["d2s_list_beta"] = np.array(R.ref_distances2similarities_list(D.copy(), P, 40, -1, 0.0, -1.0, 0, 3, pl.copy()))
    v["d2s_list_P"] = P.copy()
    P2 = np.empty(40)
    v["d2s_sigma_beta"] = np.array(R.ref_distances2similarities(D.copy(), P2, 40, -1, -1.0, 2.0, 0))
    v["d2s_sigma_P"] = P2

    # symmetrizeMatrix (tsne.cpp:1730-1828)
    row, col, val = random_csr(200, 8, seed=200, symmetric=False)
    v["sym_in_row"], v["sym_in_col"], v["sym_in_val"] = row, col, val
    sr, sc, sv = orc.symmetrize(row, col, val, lib="ref")
    v["sym_row"], v["sym_col"], v["sym_val"] = sr, sc, sv

    # evaluateErrorFft (tsne.cpp:1329-1355, nthreads = 1)
    P = random_csr(300, 8, seed=5)
    Y2 = np.random.default_rng(6).normal(scale=3, size=(300, 2))
    Y1 = np.random.default_rng(7).normal(scale=3, size=(300, 1))
    v["kl_row"], v["kl_col"], v["kl_val"], v["kl_Y2"], v["kl_Y1"] = P[0], P[1], P[2], Y2, Y1
    v["kl_out"] = np.array([orc.kl(P, Y2, 1234.5, df=1.0, lib="ref"), orc.kl(P, Y2, 1234.5, df=0.7, lib="ref"),
                            orc.kl(P, Y1, 99.0, lib="ref")])

    # exact mode: distances, dense calibration, gradient, KL (tsne.cpp:1232-1326, :1503-1530, :1832-1848)
    X, _ = mixture(60, 5, 3, seed=8)
    v["ex_X"] = X
    v["ex_dist"] = orc.squared_distances(X, lib="ref")
    v["ex_Pcal"] = orc.gaussian_perplexity_dense(X, 8.0, lib="ref")
    Pd, _ = orc.exact_symmetrize_normalize(v["ex_Pcal"])
    Ye = np.random.default_rng(9).normal(scale=2, size=(60, 2))
    v["ex_P"], v["ex_Y"] = Pd, Ye
    v["ex_grad1"] = orc.exact_gradient(Pd, Ye, df=1.0, lib="ref")[0]
    v["ex_grad2"] = orc.exact_gradient(Pd, Ye, df=0.6, lib="ref")[0]
    v["ex_kl"] = np.array([orc.exact_kl(Pd, Ye, df=1.0, lib="ref"), orc.exact_kl(Pd, Ye, df=0.6, lib="ref")])

    # Barnes-Hut (tsne.cpp:596-640, :1359-1390)
    P = random_csr(250, 6, seed=10)
    Yb = blobs(250, 2, 4, seed=11, span=10.0)
    v["bh_row"], v["bh_col"], v["bh_val"], v["bh_Y"] = P[0], P[1], P[2], Yb
    v["bh_grad"] = orc.bh_gradient(P, Yb, 0.5, lib="ref")[0]
    v["bh_err"] = np.array(orc.bh_error(P, Yb, 0.5, lib="ref"))

    # zeroMean (tsne.cpp:1851-1876)
    Z = np.random.default_rng(2).normal(loc=3.0, size=(200, 3))
    v["zm_in"] = Z
    z = Z.copy().reshape(-1)
    R.ref_zero_mean(z, 200, 3)
    v["zm_out"] = z.reshape(200, 3)

    # VP-tree kNN rows (tsne.cpp:1643-1726)
    X, _ = mixture(200, 6, 4, seed=12)
    N, K, perp = 200, 15, 5.0
    rr = np.empty(N + 1, dtype=np.uint32)
    cc = np.empty(N * K, dtype=np.uint32)
    vv = np.empty(N * K)
    R.ref_gaussian_perplexity_vptree(X.copy().reshape(-1), N, 6, perp, K, -1.0, 1, 0, np.zeros(1), 1, rr, cc, vv)
    v["vp_X"], v["vp_row"], v["vp_col"], v["vp_val"] = X, rr, cc, vv

    # TSNE::run's loop: Barnes-Hut and exact branches (tsne.cpp:389-577)
    P = random_csr(160, 6, seed=3)
    Y0 = blobs(160, 2, 4, seed=4, span=6.0, sd=1.0) * 1e-2
    with tempfile.TemporaryDirectory() as tmp:
        Yr, cr = orc.ref_tsne_run(os.path.join(tmp, "bh"), Y0, RUN_BASE, P=P)
        s = dict(RUN_BASE, no_momentum_during_exag=1, learning_rate=40.0, max_iter=9)
        Yn, cn = orc.ref_tsne_run(os.path.join(tmp, "bhn"), Y0, s, P=P)
        Pe, _ = orc.exact_symmetrize_normalize(orc.gaussian_perplexity_dense(mixture(90, 6, 3, seed=21)[0], 8.0))
        Ye0 = blobs(90, 2, 3, seed=22, span=4.0, sd=1.0) * 1e-2
        se = dict(RUN_BASE, theta=0.0, nbody_algorithm=2, df=0.5, max_iter=12)
        Ye, ce = orc.ref_tsne_run(os.path.join(tmp, "ex"), Ye0, se, Pd=Pe)
    v["run_row"], v["run_col"], v["run_val"], v["run_Y0"] = P[0], P[1], P[2], Y0
    v["run_bh_Y"], v["run_bh_costs"], v["run_bhn_Y"], v["run_bhn_costs"] = Yr, cr, Yn, cn
    v["run_ex_P"], v["run_ex_Y0"], v["run_ex_Y"], v["run_ex_costs"] = Pe, Ye0, Ye, ce

    np.savez_compressed(OUT, **v)
    print(f"wrote {OUT}: {len(v)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
