"""Python entry point of the FIt-SNE program on MI355X: same signature, defaults and file protocol as the
reference's fast_tsne.py.

`fast_tsne(X, ...)` resolves the options (learning rate, late exaggeration, PCA initialisation, search_k,
the perplexity-list and fixed-sigma switches; fast_tsne.py:177-256 of the reference), writes the
parameter/data file in the layout TSNE::load_data reads (tsne.cpp:1915-1985), runs `bin/fast_tsne`
next to this file, and reads the embedding and the loss curve back from the result file
(tsne.cpp:2025-2036).  Loss entries that were not computed (iterations whose number is not a multiple
of 50) are NaN, as in the reference.
"""
import os
import struct
import subprocess
import uuid

import numpy as np

VERSION = "1.2.1"
BINARY = os.path.join(os.path.dirname(os.path.realpath(__file__)), "bin", "fast_tsne")


def _resolve(X, theta, perplexity, map_dims, early_exag_coeff, learning_rate, start_late_exag_iter,
             stop_early_exag_iter, late_exag_coeff, initialization, seed, perplexity_list, sigma, K, search_k,
             n_trees, nbody_algo, knn_algo, load_affinities, nthreads, max_step_norm, no_momentum_during_exag):
    """The wrapper's option resolution (reference fast_tsne.py:177-256)."""
    n = X.shape[0]
    if learning_rate == "auto":
        learning_rate = max(200.0, n / early_exag_coeff)
    if start_late_exag_iter == "auto":
        start_late_exag_iter = stop_early_exag_iter if late_exag_coeff > 0 else -1
    if isinstance(initialization, str) and initialization == "pca":
        from sklearn.decomposition import PCA

        pca = PCA(n_components=map_dims, svd_solver="arpack" if X.shape[1] > map_dims else "auto",
                  random_state=seed if seed != -1 else None)
        initialization = pca.fit_transform(X)
        initialization /= np.std(initialization[:, 0])
        initialization *= 0.0001
    if perplexity_list is not None:
        perplexity = 0
    if sigma > 0 and K > 0:
        perplexity = -1
    if search_k is None:
        if perplexity > 0:
            search_k = 3 * perplexity * n_trees
        elif perplexity == 0:
            search_k = 3 * np.max(perplexity_list) * n_trees
        else:
            search_k = K * n_trees
    codes = {"Barnes-Hut": 1, "FFT": 2}
    if nbody_algo not in codes:
        raise ValueError("nbody_algo should be 'Barnes-Hut' or 'FFT'")
    knn_codes = {"vp-tree": 2, "annoy": 1}
    if knn_algo not in knn_codes:
        raise ValueError("knn_algo should be 'vp-tree' or 'annoy'")
    load = {"load": 1, "save": 2}.get(load_affinities, 0)
    return dict(learning_rate=learning_rate, start_late_exag_iter=start_late_exag_iter,
                initialization=initialization, perplexity=perplexity, search_k=search_k,
                nbody_algo=codes[nbody_algo], knn_algo=knn_codes[knn_algo], load_affinities=load,
                nthreads=0 if nthreads == -1 else nthreads, max_step_norm=-1 if max_step_norm == "none" else max_step_norm,
                no_momentum_during_exag=1 if no_momentum_during_exag else 0)


def write_data_file(path, X, theta, perplexity, perplexity_list, map_dims, max_iter, stop_early_exag_iter,
                    mom_switch_iter, momentum, final_momentum, learning_rate, max_step_norm, K, sigma, nbody_algo,
                    knn_algo, early_exag_coeff, no_momentum_during_exag, n_trees, search_k, start_late_exag_iter,
                    late_exag_coeff, nterms, intervals_per_integer, min_num_intervals, seed, df, load_affinities,
                    initialization):
    """The parameter/data file, field by field in TSNE::load_data's order (native-endian, packed)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    parts = [struct.pack("=iidd", n, d, theta, perplexity)]
    if perplexity == 0:
        parts.append(struct.pack("=i", len(perplexity_list)))
        parts.append(np.asarray(perplexity_list, dtype=np.float64).tobytes())
    parts.append(struct.pack("=iiiiddddidiidiiiididi", map_dims, max_iter, stop_early_exag_iter, mom_switch_iter,
                             momentum, final_momentum, learning_rate, max_step_norm, K, sigma, nbody_algo, knn_algo,
                             early_exag_coeff, no_momentum_during_exag, n_trees, int(search_k),
                             start_late_exag_iter, late_exag_coeff, nterms, intervals_per_integer,
                             min_num_intervals))
    parts.append(X.tobytes())
    parts.append(struct.pack("=idi", seed, df, load_affinities))
    if not isinstance(initialization, str) or initialization != "random":
        parts.append(np.ascontiguousarray(initialization, dtype=np.float64).tobytes())
    with open(path, "wb") as f:
        for p in parts:
            f.write(p)


def read_result_file(path, max_iter):
    """(Y[n, no_dims], loss[max_iter]) from the result file; uncomputed loss entries become NaN."""
    with open(path, "rb") as f:
        buf = f.read()
    n, md = struct.unpack_from("=ii", buf, 0)
    off = 8
    Y = np.frombuffer(buf, dtype=np.float64, count=n * md, offset=off).reshape(n, md).copy()
    off += 8 * n * md
    (mi,) = struct.unpack_from("=i", buf, off)
    off += 4
    loss = np.frombuffer(buf, dtype=np.float64, count=mi, offset=off).copy()
    loss[np.arange(1, mi + 1) % 50 > 0] = np.nan
    return Y, loss


def fast_tsne(X, theta=0.5, perplexity=30, map_dims=2, max_iter=750, stop_early_exag_iter=250, K=-1, sigma=-1,
              nbody_algo="FFT", knn_algo="annoy", mom_switch_iter=250, momentum=0.5, final_momentum=0.8,
              learning_rate="auto", early_exag_coeff=12, no_momentum_during_exag=False, n_trees=50, search_k=None,
              start_late_exag_iter="auto", late_exag_coeff=-1, nterms=3, intervals_per_integer=1,
              min_num_intervals=50, seed=-1, initialization="pca", load_affinities=None, perplexity_list=None, df=1,
              return_loss=False, nthreads=-1, max_step_norm=5):
    """Run t-SNE on the GPU: FFT-accelerated interpolation (nbody_algo = 'FFT'), the Barnes-Hut
    approximation (nbody_algo = 'Barnes-Hut'), or exact t-SNE with theta = 0.  Parameters and return
    value as the reference's fast_tsne (its docstring describes each option).  As in the reference, a run
    with map_dims > 2 needs nbody_algo = 'Barnes-Hut' (the FFT setting limits map_dims to 1 or 2)."""
    X = np.array(X).astype(float)
    r = _resolve(X, theta, perplexity, map_dims, early_exag_coeff, learning_rate, start_late_exag_iter,
                 stop_early_exag_iter, late_exag_coeff, initialization, seed, perplexity_list, sigma, K, search_k,
                 n_trees, nbody_algo, knn_algo, load_affinities, nthreads, max_step_norm, no_momentum_during_exag)
    tag = uuid.uuid4().hex
    infile = os.path.join(os.getcwd(), f"data_{tag}.dat")
    outfile = os.path.join(os.getcwd(), f"result_{tag}.dat")
    write_data_file(infile, X, theta, r["perplexity"], perplexity_list, map_dims, max_iter, stop_early_exag_iter,
                    mom_switch_iter, momentum, final_momentum, r["learning_rate"], r["max_step_norm"], K, sigma,
                    r["nbody_algo"], r["knn_algo"], early_exag_coeff, r["no_momentum_during_exag"], n_trees,
                    r["search_k"], r["start_late_exag_iter"], late_exag_coeff, nterms, intervals_per_integer,
                    min_num_intervals, seed, df, r["load_affinities"], r["initialization"])
    try:
        rc = subprocess.call([BINARY, VERSION, infile, outfile, str(r["nthreads"])])
        if rc != 0 or not os.path.exists(outfile):
            raise RuntimeError(f"{BINARY} failed with exit status {rc}")
        Y, loss = read_result_file(outfile, max_iter)
    finally:
        for f in (infile, outfile):
            if os.path.exists(f):
                os.remove(f)
    return (Y, loss) if return_loss else Y
