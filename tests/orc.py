"""ctypes bindings for the test oracle (oracle/liboracle.so) and the reference's own non-FFT functions
(oracle/_ref/libfitsne_ref.so, compiled from the reference sources when they are present).

Test infrastructure only: the product library never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfitsne_ref.so")
REF_RUN_SO = os.path.join(ROOT, "oracle", "_ref", "libfitsne_ref_run.so")

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
ci, cd, cu = C.c_int, C.c_double, C.c_uint

SQUARED_CAUCHY, GENERAL, SQUARED_GENERAL = 0, 1, 2


class RunParams(C.Structure):
    _fields_ = [
        ("max_iter", ci), ("stop_lying_iter", ci), ("mom_switch_iter", ci), ("start_late_exag_iter", ci),
        ("no_momentum_during_exag", ci), ("n_interpolation_points", ci), ("min_num_intervals", ci),
        ("momentum", cd), ("final_momentum", cd), ("learning_rate", cd), ("early_exag_coeff", cd),
        ("late_exag_coeff", cd), ("max_step_norm", cd), ("df", cd), ("intervals_per_integer", cd),
        ("nthreads", cu),
    ]


def _bind(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        _bind(L, "orc_kernel_1d", cd, [ci, cd, cd, cd])
        _bind(L, "orc_kernel_2d", cd, [ci, cd, cd, cd, cd, cd])
        _bind(L, "orc_fft_c2c", None, [ci, dp, dp, ci])
        _bind(L, "orc_fft_r2c_2d", None, [ci, ci, dp, dp])
        _bind(L, "orc_fft_c2r_2d", None, [ci, ci, dp, dp])
        _bind(L, "orc_dft_c2c_naive", None, [ci, dp, dp, ci])
        _bind(L, "orc_precompute_2d", None, [cd, cd, cd, cd, ci, ci, ci, dp, dp, dp, dp, dp, dp, cd])
        _bind(L, "orc_n_body_fft_2d", None, [ci, ci, dp, dp, dp, ci, ci, dp, dp, dp, dp, dp, cu])
        _bind(L, "orc_precompute_1d", None, [cd, cd, ci, ci, ci, dp, dp, dp, dp, dp, cd])
        _bind(L, "orc_n_body_fft_1d", None, [ci, ci, dp, dp, ci, ci, dp, dp, dp, dp, dp, dp])
        _bind(L, "orc_interpolate", None, [ci, ci, dp, dp, dp])
        _bind(L, "orc_minmax_2d_quirk", None, [dp, ci, C.POINTER(cd), C.POINTER(cd)])
        _bind(L, "orc_n_boxes_2d", ci, [cd, cd, cd, ci])
        _bind(L, "orc_n_boxes_1d", ci, [cd, cd, cd, ci])
        _bind(L, "orc_compute_fft_gradient", ci,
              [u32p, u32p, dp, dp, ci, ci, dp, ci, cd, ci, cu, cd, C.POINTER(cd)])
        _bind(L, "orc_attractive", None, [u32p, u32p, dp, dp, ci, ci, cd, cu, dp])
        _bind(L, "orc_exact_repulsion", None, [dp, ci, ci, cd, dp, C.POINTER(cd)])
        _bind(L, "orc_zero_mean", None, [dp, ci, ci])
        _bind(L, "orc_update", None, [dp, dp, dp, dp, ci, ci, cd, cd, cd, ci])
        _bind(L, "orc_evaluate_error_fft", cd, [u32p, u32p, dp, dp, ci, ci, cd, cd])
        _bind(L, "orc_run_fft", ci, [u32p, u32p, dp, ci, ci, dp, C.POINTER(RunParams), dp])
        _bind(L, "orc_run_bh", ci, [u32p, u32p, dp, ci, ci, dp, C.POINTER(RunParams), cd, dp])
        _bind(L, "orc_run_exact", ci, [dp, ci, ci, dp, C.POINTER(RunParams), dp])
        _bind(L, "orc_distances2similarities", cd, [dp, dp, ci, ci, cd, cd, ci])
        _bind(L, "orc_d2s_bisect", cd, [dp, dp, ci, cd, cd, ci, C.POINTER(ci), C.POINTER(cd)])
        _bind(L, "orc_distances2similarities_list", cd, [dp, dp, ci, ci, cd, cd, ci, ci, dp])
        _bind(L, "orc_knn_exact", None, [dp, ci, ci, ci, i32p, dp, cu])
        _bind(L, "orc_rows_from_knn", None, [i32p, dp, ci, ci, cd, cd, ci, dp, u32p, u32p, dp])
        _bind(L, "orc_symmetrize", ci, [u32p, u32p, dp, ci, u32p, u32p, dp])
        _bind(L, "orc_squared_distances", None, [dp, ci, ci, dp])
        _bind(L, "orc_gaussian_perplexity_dense", None, [dp, ci, ci, cd, cd, ci, dp, dp, cu])
        _bind(L, "orc_exact_symmetrize_normalize", cd, [dp, ci])
        _bind(L, "orc_exact_gradient", None, [dp, dp, ci, ci, cd, dp, C.POINTER(cd)])
        _bind(L, "orc_evaluate_error_exact", cd, [dp, dp, ci, ci, cd])
        _bind(L, "orc_bh_gradient", None, [u32p, u32p, dp, dp, ci, ci, cd, dp, C.POINTER(cd)])
        _bind(L, "orc_bh_error", cd, [u32p, u32p, dp, dp, ci, ci, cd])
        _oracle = L
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        _bind(L, "ref_interpolate", None, [ci, ci, dp, dp, dp])
        _bind(L, "ref_kernel_1d", cd, [ci, cd, cd, cd])
        _bind(L, "ref_kernel_2d", cd, [ci, cd, cd, cd, cd, cd])
        _bind(L, "ref_distances2similarities", cd, [dp, dp, ci, ci, cd, cd, ci])
        _bind(L, "ref_distances2similarities_list", cd, [dp, dp, ci, ci, cd, cd, ci, ci, dp])
        _bind(L, "ref_symmetrize", ci, [u32p, u32p, dp, ci, u32p, u32p, dp])
        _bind(L, "ref_evaluate_error_fft", cd, [u32p, u32p, dp, dp, ci, ci, cd, cd])
        _bind(L, "ref_exact_gradient", None, [dp, dp, ci, ci, dp, cd])
        _bind(L, "ref_evaluate_error_exact", cd, [dp, dp, ci, ci, cd])
        _bind(L, "ref_zero_mean", None, [dp, ci, ci])
        _bind(L, "ref_squared_distances", None, [dp, ci, ci, dp])
        _bind(L, "ref_gaussian_perplexity_dense", None, [dp, ci, ci, dp, cd, cd, ci, dp])
        _bind(L, "ref_bh_gradient", None, [u32p, u32p, dp, dp, ci, ci, cd, dp])
        _bind(L, "ref_bh_error", cd, [u32p, u32p, dp, dp, ci, ci, cd])
        _bind(L, "ref_gaussian_perplexity_vptree", None,
              [dp, ci, ci, cd, ci, cd, cu, ci, dp, cu, u32p, u32p, dp])
        _bind(L, "ref_load_data", ci, [C.c_char_p, i32p, dp, C.c_void_p, C.c_void_p, C.c_void_p])
        _bind(L, "ref_save_data", None, [C.c_char_p, dp, dp, ci, ci, ci])
        _ref = L
    return _ref


_ref_run = None


def have_ref_run():
    return os.path.exists(REF_RUN_SO)


def ref_run():
    """The reference's TSNE::run (oracle/_ref/libfitsne_ref_run.so).  FFTW's entry points are unresolved in
    that library, so it is opened with lazy binding; only its Barnes-Hut and exact branches are driven.
    It reads its affinities from the current directory (load_affinities = 1): call it from a directory of
    its own (tests/ref_run_child.py)."""
    global _ref_run
    if _ref_run is None:
        L = C.CDLL(REF_RUN_SO, mode=os.RTLD_LAZY)
        _bind(L, "ref_tsne_run", ci, [dp, ci, ci, dp, ci, cd, cd, ci, ci, ci, cd, cd, cd, ci, cd, dp, ci, ci, cd,
                                      cd, cd])
        _ref_run = L
    return _ref_run


def ref_tsne_run(workdir, Y0, settings, P=None, Pd=None):
    """The reference's TSNE::run with the given affinities, in a child process whose working directory is
    `workdir` (tests/ref_run_child.py): the CSR P is written to P_row.dat / P_col.dat / P_val.dat, a dense
    Pd to P.dat, where its load_affinities = 1 branch reads them (tsne.cpp:172-276).  `settings` holds
    perplexity, theta, max_iter, stop_lying_iter, mom_switch_iter, momentum, final_momentum,
    learning_rate, nbody_algorithm, early_exag_coeff, no_momentum_during_exag, start_late_exag_iter,
    late_exag_coeff, df and max_step_norm.  Returns (Y, costs)."""
    import subprocess
    import sys
    os.makedirs(workdir, exist_ok=True)
    if P is not None:
        row, col, val = P
        np.asarray(row, dtype=np.uint32).tofile(os.path.join(workdir, "P_row.dat"))
        np.asarray(col, dtype=np.uint32).tofile(os.path.join(workdir, "P_col.dat"))
        np.asarray(val, dtype=np.float64).tofile(os.path.join(workdir, "P_val.dat"))
    if Pd is not None:
        np.ascontiguousarray(Pd, dtype=np.float64).tofile(os.path.join(workdir, "P.dat"))
    np.savez(os.path.join(workdir, "run.npz"), Y0=np.asarray(Y0, dtype=np.float64), **settings)
    child = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_run_child.py")
    r = subprocess.run([sys.executable, child], cwd=workdir, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError("reference run failed: " + r.stdout[-2000:] + r.stderr[-2000:])
    out = np.load(os.path.join(workdir, "out.npz"))
    if int(out["rc"]) != 0:
        raise RuntimeError(f"TSNE::run returned {int(out['rc'])}")
    return out["Y"], out["costs"]


# ---------------------------------------------------------------- convenience wrappers

def fft_c2c(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    oracle().orc_fft_c2c(x.size, x.view(np.float64), out.view(np.float64), int(inverse))
    return out


def dft_naive(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    oracle().orc_dft_c2c_naive(x.size, x.view(np.float64), out.view(np.float64), int(inverse))
    return out


def fft_r2c_2d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n0, n1 = a.shape
    out = np.empty((n0, n1 // 2 + 1), dtype=np.complex128)
    oracle().orc_fft_r2c_2d(n0, n1, a, out.view(np.float64))
    return out


def fft_c2r_2d(z, n1):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    n0 = z.shape[0]
    out = np.empty((n0, n1), dtype=np.float64)
    oracle().orc_fft_c2r_2d(n0, n1, z.view(np.float64), out)
    return out


def interpolate(p, y_in_box, spacings, lib="oracle"):
    y = np.ascontiguousarray(y_in_box, dtype=np.float64)
    s = np.ascontiguousarray(spacings, dtype=np.float64)
    out = np.empty(p * y.size)
    if lib == "oracle":
        oracle().orc_interpolate(p, y.size, y, s, out)
    else:
        ref().ref_interpolate(p, y.size, y, s, out)
    return out.reshape(p, y.size)


def precompute_2d(xmax, xmin, nb, p, kernel, df=1.0):
    n1d = nb * p
    M = 2 * n1d
    lo = np.empty(2 * nb * nb)
    hi = np.empty(2 * nb * nb)
    sp = np.empty(p)
    yt = np.empty(n1d)
    xt = np.empty(n1d)
    kt = np.empty((M, M // 2 + 1), dtype=np.complex128)
    oracle().orc_precompute_2d(xmax, xmin, xmax, xmin, nb, p, kernel, lo, hi, sp, yt, xt, kt.view(np.float64), df)
    return dict(lo=lo, hi=hi, sp=sp, yt=yt, xt=xt, kt=kt)


def n_body_fft_2d(xs, ys, charges, nb, p, geo, nthreads=1):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    N, nt = ch.shape
    pot = np.zeros((N, nt))
    oracle().orc_n_body_fft_2d(N, nt, xs, ys, ch, nb, p, geo["lo"], geo["hi"], geo["sp"],
                               geo["kt"].view(np.float64), pot, nthreads)
    return pot


def precompute_1d(ymin, ymax, nb, p, kernel, df=1.0):
    lo = np.empty(nb)
    hi = np.empty(nb)
    sp = np.empty(p)
    yt = np.empty(nb * p)
    kv = np.empty(2 * nb * p, dtype=np.complex128)
    oracle().orc_precompute_1d(ymin, ymax, nb, p, kernel, lo, hi, sp, yt, kv.view(np.float64), df)
    return dict(lo=lo, hi=hi, sp=sp, yt=yt, kv=kv)


def n_body_fft_1d(Y, charges, nb, p, geo):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    N, nt = ch.shape
    pot = np.zeros((N, nt))
    oracle().orc_n_body_fft_1d(N, nt, Y, ch, nb, p, geo["lo"], geo["hi"], geo["sp"], geo["yt"],
                               geo["kv"].view(np.float64), pot)
    return pot


def gradient(P, Y, p=3, ipi=1.0, min_num_intervals=50, nthreads=1, df=1.0):
    """P = (row_P u32, col_P u32, val_P f64).  Returns (dC [N, D], sum_Q)."""
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape if Y.ndim == 2 else (Y.size, 1)
    dC = np.empty(N * D)
    sq = C.c_double()
    rc = oracle().orc_compute_fft_gradient(row, col, val, Y.reshape(-1), N, D, dC, p, ipi, min_num_intervals,
                                           nthreads, df, C.byref(sq))
    assert rc == 0
    return dC.reshape(N, D) if D > 1 else dC.reshape(N, 1), sq.value


def attractive(P, Y, df=1.0, nthreads=1):
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    out = np.empty(N * D)
    oracle().orc_attractive(row, col, val, Y.reshape(-1), N, D, df, nthreads, out)
    return out.reshape(N, D)


def exact_repulsion(Y, df=1.0):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    out = np.empty(N * D)
    sq = C.c_double()
    oracle().orc_exact_repulsion(Y.reshape(-1), N, D, df, out, C.byref(sq))
    return out.reshape(N, D), sq.value


def kl(P, Y, sum_Q, df=1.0, lib="oracle"):
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    if lib == "oracle":
        return oracle().orc_evaluate_error_fft(row, col, val, Y.reshape(-1), N, D, sum_Q, df)
    return ref().ref_evaluate_error_fft(row, col, val, Y.reshape(-1).copy(), N, D, sum_Q, df)


def minmax_quirk(Y):
    a, b = C.c_double(), C.c_double()
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    oracle().orc_minmax_2d_quirk(Y.reshape(-1), Y.shape[0], C.byref(a), C.byref(b))
    return a.value, b.value


def knn_exact(X, K, nthreads=8):
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    idx = np.empty(N * K, dtype=np.int32)
    dist = np.empty(N * K)
    oracle().orc_knn_exact(X.reshape(-1), N, D, K, idx, dist, nthreads)
    return idx.reshape(N, K), dist.reshape(N, K)


def rows_from_knn(idx, dist, perplexity, sigma=-1.0, plist=None):
    N, K = idx.shape
    row = np.empty(N + 1, dtype=np.uint32)
    col = np.empty(N * K, dtype=np.uint32)
    val = np.empty(N * K)
    pl = np.ascontiguousarray(plist if plist is not None else np.zeros(1), dtype=np.float64)
    oracle().orc_rows_from_knn(np.ascontiguousarray(idx, dtype=np.int32).reshape(-1),
                               np.ascontiguousarray(dist).reshape(-1), N, K, perplexity, sigma,
                               0 if plist is None else len(plist), pl, row, col, val)
    return row, col, val


def symmetrize(row, col, val, lib="oracle"):
    N = row.size - 1
    nnz = int(row[-1])
    sr = np.empty(N + 1, dtype=np.uint32)
    sc = np.empty(2 * nnz, dtype=np.uint32)
    sv = np.empty(2 * nnz)
    fn = oracle().orc_symmetrize if lib == "oracle" else ref().ref_symmetrize
    n = fn(row, col, val, N, sr, sc, sv)
    return sr, sc[:n].copy(), sv[:n].copy()


def normalize(row, col, val):
    """tsne.cpp:327-330: val_P /= sum(val_P) after symmetrisation."""
    s = 0.0
    for v in val:  # reference order, serial
        s += v
    return row, col, val / s


def run_params(**kw):
    return RunParams(
        max_iter=kw.get("max_iter", 1000), stop_lying_iter=kw.get("stop_lying_iter", 250),
        mom_switch_iter=kw.get("mom_switch_iter", 250), start_late_exag_iter=kw.get("start_late_exag_iter", -1),
        no_momentum_during_exag=kw.get("no_momentum_during_exag", 0),
        n_interpolation_points=kw.get("nterms", 3), min_num_intervals=kw.get("min_num_intervals", 50),
        momentum=kw.get("momentum", 0.5), final_momentum=kw.get("final_momentum", 0.8),
        learning_rate=kw.get("learning_rate", 200.0), early_exag_coeff=kw.get("early_exag_coeff", 12.0),
        late_exag_coeff=kw.get("late_exag_coeff", -1.0), max_step_norm=kw.get("max_step_norm", 5.0),
        df=kw.get("df", 1.0), intervals_per_integer=kw.get("intervals_per_integer", 1.0),
        nthreads=kw.get("nthreads", 1))


def run_fft(P, Y0, **kw):
    """The oracle's TSNE::run loop with the FFT gradient (orc_run_fft)."""
    row, col, val = P
    Y = np.ascontiguousarray(Y0, dtype=np.float64).copy()
    N, D = Y.shape
    prm = run_params(**kw)
    costs = np.zeros(prm.max_iter)
    rc = oracle().orc_run_fft(row, col, val, N, D, Y.reshape(-1), C.byref(prm), costs)
    assert rc == 0
    return Y, costs


def run_bh(P, Y0, theta=0.5, **kw):
    """The same loop with the Barnes-Hut gradient and KL (orc_run_bh)."""
    row, col, val = P
    Y = np.ascontiguousarray(Y0, dtype=np.float64).copy()
    N, D = Y.shape
    prm = run_params(**kw)
    costs = np.zeros(prm.max_iter)
    rc = oracle().orc_run_bh(row, col, np.ascontiguousarray(val, dtype=np.float64), N, D, Y.reshape(-1),
                             C.byref(prm), theta, costs)
    assert rc == 0
    return Y, costs


def run_exact(Pd, Y0, **kw):
    """The same loop with the exact gradient over a dense P (orc_run_exact)."""
    Y = np.ascontiguousarray(Y0, dtype=np.float64).copy()
    N, D = Y.shape
    prm = run_params(**kw)
    costs = np.zeros(prm.max_iter)
    rc = oracle().orc_run_exact(np.ascontiguousarray(Pd, dtype=np.float64).reshape(-1), N, D, Y.reshape(-1),
                                C.byref(prm), costs)
    assert rc == 0
    return Y, costs


# ---------------------------------------------------------------- exact t-SNE (dense P)

def squared_distances(X, lib="oracle"):
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    DD = np.empty(N * N)
    if lib == "oracle":
        oracle().orc_squared_distances(X.reshape(-1), N, D, DD)
    else:
        ref().ref_squared_distances(X.reshape(-1).copy(), N, D, DD)
    return DD.reshape(N, N)


def gaussian_perplexity_dense(X, perplexity, sigma=-1.0, plist=None, lib="oracle", nthreads=8):
    """Row-calibrated dense P before symmetrisation (tsne.cpp:1503-1530)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    P = np.empty(N * N)
    L = 0 if plist is None else len(plist)
    pl = np.ascontiguousarray(plist if plist is not None else np.zeros(1), dtype=np.float64)
    if lib == "oracle":
        oracle().orc_gaussian_perplexity_dense(X.reshape(-1), N, D, perplexity, sigma, L, pl, P, nthreads)
    else:
        ref().ref_gaussian_perplexity_dense(X.reshape(-1).copy(), N, D, P, perplexity, sigma, L, pl.copy())
    return P.reshape(N, N)


def exact_symmetrize_normalize(P):
    P = np.ascontiguousarray(P, dtype=np.float64).copy()
    s = oracle().orc_exact_symmetrize_normalize(P.reshape(-1), P.shape[0])
    return P, s


def exact_P(X, perplexity, sigma=-1.0, plist=None):
    """The exact mode's P (tsne.cpp:171-229): calibrated rows, P + P^T, divided by the total."""
    return exact_symmetrize_normalize(gaussian_perplexity_dense(X, perplexity, sigma, plist))


def exact_gradient(P, Y, df=1.0, lib="oracle"):
    P = np.ascontiguousarray(P, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    dC = np.empty(N * D)
    if lib == "oracle":
        sq = C.c_double()
        oracle().orc_exact_gradient(P.reshape(-1), Y.reshape(-1), N, D, df, dC, C.byref(sq))
        return dC.reshape(N, D), sq.value
    ref().ref_exact_gradient(P.reshape(-1).copy(), Y.reshape(-1).copy(), N, D, dC, df)
    return dC.reshape(N, D), None


def exact_kl(P, Y, df=1.0, lib="oracle"):
    P = np.ascontiguousarray(P, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    if lib == "oracle":
        return oracle().orc_evaluate_error_exact(P.reshape(-1), Y.reshape(-1), N, D, df)
    return ref().ref_evaluate_error_exact(P.reshape(-1).copy(), Y.reshape(-1).copy(), N, D, df)


# ---------------------------------------------------------------- Barnes-Hut (SPTree)

def _u32csr(P):
    row, col, val = P
    return (np.ascontiguousarray(row, dtype=np.uint32), np.ascontiguousarray(col, dtype=np.uint32),
            np.ascontiguousarray(val, dtype=np.float64))


def bh_gradient(P, Y, theta, lib="oracle"):
    row, col, val = _u32csr(P)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    dC = np.empty(N * D)
    if lib == "oracle":
        sq = C.c_double()
        oracle().orc_bh_gradient(row, col, val, Y.reshape(-1), N, D, theta, dC, C.byref(sq))
        return dC.reshape(N, D), sq.value
    ref().ref_bh_gradient(row.copy(), col.copy(), val.copy(), Y.reshape(-1).copy(), N, D, theta, dC)
    return dC.reshape(N, D), None


def bh_error(P, Y, theta, lib="oracle"):
    row, col, val = _u32csr(P)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N, D = Y.shape
    if lib == "oracle":
        return oracle().orc_bh_error(row, col, val, Y.reshape(-1), N, D, theta)
    return ref().ref_bh_error(row.copy(), col.copy(), val.copy(), Y.reshape(-1).copy(), N, D, theta)
