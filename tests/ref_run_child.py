"""Runs the reference's own TSNE::run (oracle/_ref/libfitsne_ref_run.so) once, in the current directory.

Test infrastructure, started as a child process by tests/test_oracle_run.py (and by
tests/golden/make_ref_golden.py) with its own working directory, because the reference reads its
affinities from P.dat or P_row.dat / P_col.dat / P_val.dat there (load_affinities = 1,
tsne.cpp:172-276).  Reads run.npz (Y0 and the scalar settings), writes out.npz (Y, costs, rc).
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import orc  # noqa: E402


def main():
    a = np.load("run.npz")
    Y = np.ascontiguousarray(a["Y0"], dtype=np.float64).copy()
    N, nd = Y.shape
    s = {k: a[k].item() for k in a.files if k != "Y0"}
    X = np.zeros(N * 2)  # the loaded-affinities branch only centres and scales X (tsne.cpp:149-162)
    costs = np.zeros(int(s["max_iter"]))
    rc = orc.ref_run().ref_tsne_run(
        X, N, 2, Y.reshape(-1), nd, float(s["perplexity"]), float(s["theta"]), int(s["max_iter"]),
        int(s["stop_lying_iter"]), int(s["mom_switch_iter"]), float(s["momentum"]), float(s["final_momentum"]),
        float(s["learning_rate"]), int(s["nbody_algorithm"]), float(s["early_exag_coeff"]), costs,
        int(s["no_momentum_during_exag"]), int(s["start_late_exag_iter"]), float(s["late_exag_coeff"]),
        float(s["df"]), float(s["max_step_norm"]))
    np.savez("out.npz", Y=Y, costs=costs, rc=rc)


if __name__ == "__main__":
    main()
