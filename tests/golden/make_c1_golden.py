"""Writes tests/golden/c1_oracle_kl.json: the oracle's loss curve for the C1 configuration.

C1: N = 10,000 points in D = 50 (10 equal-weight Gaussians, centres N(0, 4^2 I), sd 1, fit-sne_amd/synth.py
c1_data), perplexity 30 with K = 90 exact neighbours, symmetric P, 2D embedding from a PCA initialisation
scaled to sd 1e-4 (c1_init below), 1000 iterations, early exaggeration 12 until iteration 250, learning
rate max(200, N/12), momentum 0.5 -> 0.8 at 250, max_step_norm 5, n_interp 3, min_num_intervals 50, df 1.
The run is the oracle's restatement of TSNE::run (tests/orc.py run_fft) with P built by the oracle's
restatement of computeGaussianPerplexity + symmetrizeMatrix.  It takes a few minutes on 8 cores, so the
GPU end-to-end test compares against this file instead of rerunning it.

Usage (from the repository root): python tests/golden/make_c1_golden.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fit-sne_amd"))

import orc  # noqa: E402
import synth  # noqa: E402


def c1_init(X):
    """PCA (two leading right singular vectors of the centred data), scaled so that column 0 has sd 1e-4;
    the sign of each component is fixed by making its largest-magnitude loading positive."""
    Xc = X - X.mean(axis=0)
    _, _, vt = np.linalg.svd(Xc, full_matrices=False)
    V = vt[:2].T
    V *= np.sign(V[np.argmax(np.abs(V), axis=0), [0, 1]])
    Y0 = Xc @ V
    Y0 /= np.std(Y0[:, 0])
    return np.ascontiguousarray(Y0 * 1e-4)


def c1_affinities(X, perplexity=30.0):
    N, D = X.shape
    Xc = np.ascontiguousarray(X, dtype=np.float64).copy()
    orc.oracle().orc_zero_mean(Xc.reshape(-1), N, D)
    Xc /= np.abs(Xc).max()
    idx, dist = orc.knn_exact(Xc, int(3 * perplexity), nthreads=os.cpu_count() or 1)
    row, col, val = orc.rows_from_knn(idx, dist, perplexity)
    return orc.normalize(*orc.symmetrize(row, col, val))


def main():
    X, _ = synth.c1_data()
    N = X.shape[0]
    P = c1_affinities(X)
    Y0 = c1_init(X)
    t0 = time.time()
    _, costs = orc.run_fft(P, Y0, max_iter=1000, learning_rate=max(200.0, N / 12.0), nthreads=os.cpu_count() or 1)
    out = {"config": "C1 N=10000 D=50 perplexity 30, 1000 iterations, early exaggeration 12 until 250",
           "nnz": int(P[0][-1]), "iterations": [i + 1 for i in range(49, 1000, 50)],
           "kl": [float(costs[i]) for i in range(49, 1000, 50)], "seconds": time.time() - t0}
    with open(os.path.join(HERE, "c1_oracle_kl.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
