"""ctypes binding of libfitsne.so (include/fitsne_abi.h) for tests, benchmarks and the Python wrapper.

The GPU library is the only implementation behind these calls: if lib/libfitsne.so is missing or no
GPU is visible, calls raise instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfitsne.so")

OK, E_ARG, E_ALLOC, E_HIP, E_UNSUPPORTED, E_NODEVICE = 0, -1, -2, -3, -4, -5
KERNEL_SQUARED_CAUCHY, KERNEL_GENERAL, KERNEL_SQUARED_GENERAL = 0, 1, 2
FP64_PLAIN, FP32, FP64_EXACT, FP64_BH = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ci, _cd, _cu, _vp = C.c_int, C.c_double, C.c_uint, C.c_void_p


class FitsneError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fitsne error {code}: {msg}")
        self.code = code


class StepParams(C.Structure):
    _fields_ = [("momentum", _cd), ("learning_rate", _cd), ("max_step_norm", _cd),
                ("plain_descent", _ci), ("reserved", _ci)]


# every symbol include/fitsne_abi.h declares, with its ctypes signature
SIGNATURES = {
    "fitsne_last_error": (C.c_char_p, []),
    "fitsne_device_count": (_ci, []),
    "fitsne_version": (C.c_char_p, []),
    "fitsne_compute_fft_gradient": (_ci, [_u32, _u32, _dp, _dp, _ci, _ci, _dp, _ci, _cd, _ci, _cu, _cd, C.POINTER(_cd)]),
    "fitsne_compute_fft_gradient_i64": (_ci, [_i64, _i32, _dp, _dp, _ci, _ci, _dp, _ci, _cd, _ci, _cu, _cd, C.POINTER(_cd)]),
    "fitsne_precompute_2d": (_ci, [_cd, _cd, _cd, _cd, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp, _cd]),
    "fitsne_n_body_fft_2d": (_ci, [_ci, _ci, _dp, _dp, _dp, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _cu]),
    "fitsne_precompute_1d": (_ci, [_cd, _cd, _ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _cd]),
    "fitsne_n_body_fft_1d": (_ci, [_ci, _ci, _dp, _dp, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp]),
    "fitsne_interpolate": (_ci, [_ci, _ci, _dp, _dp, _dp]),
    "fitsne_evaluate_error_fft": (_ci, [_u32, _u32, _dp, _dp, _ci, _ci, _cd, _cd, C.POINTER(_cd)]),
    "fitsne_evaluate_error_fft_noalloc": (_ci, [_u32, _u32, _dp, _dp, _ci, _ci, _cd, _cd, C.POINTER(_cd)]),
    "fitsne_distances2similarities_batch": (_ci, [_dp, _dp, _ci, _ci, _cd, _cd, _ci, _vp, _dp, _vp]),
    "fitsne_knn_exact": (_ci, [_dp, _ci, _ci, _ci, _i32, _dp]),
    "fitsne_knn_mfma": (_ci, [_dp, _ci, _ci, _ci, _i32, _dp, C.POINTER(C.c_longlong)]),
    "fitsne_symmetrize": (_ci, [_u32, _u32, _dp, _ci, _u32, _u32, _dp, C.POINTER(C.c_longlong)]),
    "fitsne_set_graph": (_ci, [_vp, _ci]),
    "fitsne_set_overlap": (_ci, [_vp, _ci]),
    "fitsne_autotune": (_ci, [_vp, C.POINTER(_ci)]),
    "fitsne_autotune_stats": (_ci, [_vp, C.POINTER(_ci), C.POINTER(_ci)]),
    "fitsne_set_attractive_rows": (_ci, [_vp, _ci]),
    "fitsne_set_spread_points": (_ci, [_vp, _ci]),
    "fitsne_batch_create": (_ci, [C.POINTER(_vp), C.POINTER(_vp), _ci]),
    "fitsne_batch_step": (_ci, [_vp, C.POINTER(StepParams)]),
    "fitsne_batch_destroy": (_ci, [_vp]),
    "fitsne_beta_model_create": (_ci, [C.POINTER(_vp), _ci, C.c_ulonglong]),
    "fitsne_beta_model_destroy": (_ci, [_vp]),
    "fitsne_beta_model_fit": (_ci, [_vp, _dp, _dp, _ci, _ci, _ci, _cd, C.c_ulonglong, C.POINTER(_cd)]),
    "fitsne_beta_model_predict": (_ci, [_vp, _dp, _ci, _dp]),
    "fitsne_beta_model_gradient": (_ci, [_vp, _dp, _dp, _ci, _f32, C.POINTER(_cd)]),
    "fitsne_beta_model_params": (_ci, [_vp, _f32]),
    "fitsne_beta_model_set_params": (_ci, [_vp, _f32]),
    "fitsne_ctx_create": (_ci, [C.POINTER(_vp), _ci, _ci, _ci, _ci, _cd, _ci, _cd]),
    "fitsne_ctx_destroy": (_ci, [_vp]),
    "fitsne_set_P": (_ci, [_vp, _ci, _u32, _u32, _dp]),
    "fitsne_set_P_i64": (_ci, [_vp, _ci, _i64, _i32, _dp]),
    "fitsne_exact_affinities": (_ci, [_vp, _dp, _ci, _ci, _cd, _cd, _ci, _vp, _vp, C.POINTER(_cd)]),
    "fitsne_set_P_dense": (_ci, [_vp, _ci, _dp]),
    "fitsne_set_theta": (_ci, [_vp, _cd]),
    "fitsne_set_exaggeration": (_ci, [_vp, _cd]),
    "fitsne_set_reorder": (_ci, [_vp, _ci]),
    "fitsne_reorder": (_ci, [_vp]),
    "fitsne_set_Y": (_ci, [_vp, _dp]),
    "fitsne_get_Y": (_ci, [_vp, _dp]),
    "fitsne_gradient": (_ci, [_vp, _vp, C.POINTER(_cd)]),
    "fitsne_step": (_ci, [_vp, C.POINTER(StepParams), _vp]),
    "fitsne_kl": (_ci, [_vp, C.POINTER(_cd)]),
    "fitsne_sync": (_ci, [_vp]),
    "fitsne_set_timing": (_ci, [_vp, _ci]),
    "fitsne_kernel_stats": (_ci, [_vp, _ci, C.c_char_p, _vp, _vp]),
    "fitsne_last_geometry": (_ci, [_vp, _dp]),
    "fitsne_debug_read": (C.c_longlong, [_vp, C.c_char_p, _vp, C.c_longlong]),
}

_lib = None


def lib():
    """Load libfitsne.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FitsneError(E_NODEVICE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise FitsneError(rc, lib().fitsne_last_error().decode())
    return rc


def device_count():
    return lib().fitsne_device_count()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------- mirror tier

def compute_fft_gradient(P, Y, n_interpolation_points=3, intervals_per_integer=1.0, min_num_intervals=50,
                         df=1.0, nthreads=0):
    """Mirror of TSNE::computeFftGradient* (FP64 plain GPU path).  P = (row_P, col_P, val_P)."""
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N = Y.shape[0]
    D = 1 if Y.ndim == 1 else Y.shape[1]
    dC = np.empty(N * D)
    sq = _cd()
    if row.dtype == np.int64:
        check(lib().fitsne_compute_fft_gradient_i64(row, np.ascontiguousarray(col, dtype=np.int32), val,
                                                    Y.reshape(-1), N, D, dC, n_interpolation_points,
                                                    intervals_per_integer, min_num_intervals, nthreads, df,
                                                    C.byref(sq)))
    else:
        check(lib().fitsne_compute_fft_gradient(row, col, val, Y.reshape(-1), N, D, dC, n_interpolation_points,
                                                intervals_per_integer, min_num_intervals, nthreads, df,
                                                C.byref(sq)))
    return dC.reshape(N, D), sq.value


def precompute_2d(xmax, xmin, nb, p, kernel, df=1.0):
    n1d = nb * p
    M = 2 * n1d
    lo = np.empty(2 * nb * nb)
    hi = np.empty(2 * nb * nb)
    sp = np.empty(p)
    yt = np.empty(n1d)
    xt = np.empty(n1d)
    kt = np.empty(2 * M * (M // 2 + 1))
    check(lib().fitsne_precompute_2d(xmax, xmin, xmax, xmin, nb, p, kernel, lo, hi, sp, yt, xt, kt, df))
    return dict(lo=lo, hi=hi, sp=sp, yt=yt, xt=xt, kt=kt)


def n_body_fft_2d(xs, ys, charges, nb, p, geo, potential=None):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    N, nt = ch.shape
    pot = np.zeros((N, nt)) if potential is None else np.ascontiguousarray(potential, dtype=np.float64).copy()
    check(lib().fitsne_n_body_fft_2d(N, nt, xs, ys, ch.reshape(-1), nb, p, geo["lo"], geo["hi"], geo["sp"],
                                     geo["kt"], pot.reshape(-1), 0))
    return pot


def precompute_1d(ymin, ymax, nb, p, kernel, df=1.0):
    lo = np.empty(nb)
    hi = np.empty(nb)
    sp = np.empty(p)
    yt = np.empty(nb * p)
    kv = np.empty(4 * nb * p)
    check(lib().fitsne_precompute_1d(ymin, ymax, nb, p, kernel, lo, hi, sp, yt, kv, df))
    return dict(lo=lo, hi=hi, sp=sp, yt=yt, kv=kv)


def n_body_fft_1d(Y, charges, nb, p, geo, potential=None):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    N, nt = ch.shape
    pot = np.zeros((N, nt)) if potential is None else np.ascontiguousarray(potential, dtype=np.float64).copy()
    check(lib().fitsne_n_body_fft_1d(N, nt, Y, ch.reshape(-1), nb, p, geo["lo"], geo["hi"], geo["sp"], geo["yt"],
                                     geo["kv"], pot.reshape(-1)))
    return pot


def interpolate(p, y_in_box, spacings):
    y = np.ascontiguousarray(y_in_box, dtype=np.float64)
    out = np.empty(p * y.size)
    check(lib().fitsne_interpolate(p, y.size, y, np.ascontiguousarray(spacings, dtype=np.float64), out))
    return out.reshape(p, y.size)


def evaluate_error_fft(P, Y, sum_Q, df=1.0):
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N = Y.shape[0]
    D = 1 if Y.ndim == 1 else Y.shape[1]
    out = _cd()
    check(lib().fitsne_evaluate_error_fft(row, col, val, Y.reshape(-1), N, D, sum_Q, df, C.byref(out)))
    return out.value


def evaluate_error_fft_noalloc(P, Y, sum_Q, df=1.0):
    """evaluate_error_fft through the entry point that allocates no device memory (fixed 32 MiB scratch)."""
    row, col, val = P
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    N = Y.shape[0]
    D = 1 if Y.ndim == 1 else Y.shape[1]
    out = _cd()
    check(lib().fitsne_evaluate_error_fft_noalloc(row, col, val, Y.reshape(-1), N, D, sum_Q, df, C.byref(out)))
    return out.value


def distances2similarities(Dist, perplexity, sigma=-1.0, squared=False, beta0=None):
    Dist = np.ascontiguousarray(Dist, dtype=np.float64)
    n, K = Dist.shape
    P = np.empty((n, K))
    beta = np.empty(n)
    iters = np.empty(n, dtype=np.int32)
    b0 = None if beta0 is None else np.ascontiguousarray(beta0, dtype=np.float64)
    check(lib().fitsne_distances2similarities_batch(Dist.reshape(-1), P.reshape(-1), n, K, perplexity, sigma,
                                                    int(squared), _ptr(b0), beta, _ptr(iters)))
    return P, beta, iters


def knn_exact(X, K):
    """Exact K nearest other rows of X (FP64), ascending by (distance, index): (idx[N,K], dist[N,K])."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    idx = np.empty((N, K), dtype=np.int32)
    dist = np.empty((N, K))
    check(lib().fitsne_knn_exact(X.reshape(-1), N, D, K, idx.reshape(-1), dist.reshape(-1)))
    return idx, dist


def knn_mfma(X, K):
    """The same neighbours as knn_exact for large N (FP32 MFMA candidates, FP64 re-rank): (idx, dist,
    number of queries recomputed by the FP64 kernel)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    idx = np.empty((N, K), dtype=np.int32)
    dist = np.empty((N, K))
    redo = C.c_longlong(0)
    check(lib().fitsne_knn_mfma(X.reshape(-1), N, D, K, idx.reshape(-1), dist.reshape(-1), C.byref(redo)))
    return idx, dist, redo.value


def symmetrize(row, col, val):
    """symmetrizeMatrix on the GPU: returns (sym_row, sym_col, sym_val) before normalisation."""
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    N = row.size - 1
    nnz = int(row[-1])
    sr = np.empty(N + 1, dtype=np.uint32)
    sc = np.empty(max(2 * nnz, 1), dtype=np.uint32)
    sv = np.empty(max(2 * nnz, 1))
    m = C.c_longlong()
    check(lib().fitsne_symmetrize(row, col, val, N, sr, sc, sv, C.byref(m)))
    return sr, sc[: m.value].copy(), sv[: m.value].copy()


class Batch:
    """One step of several FP32 contexts (independent embeddings) as one graph launch."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (_vp * len(self.contexts))(*[c.h for c in self.contexts])
        h = _vp()
        check(lib().fitsne_batch_create(C.byref(h), arr, len(self.contexts)))
        self.h = h

    def step(self, params):
        """params: one (momentum, learning_rate, max_step_norm) tuple per context, or one for all."""
        if isinstance(params, tuple):
            params = [params] * len(self.contexts)
        arr = (StepParams * len(self.contexts))(*[StepParams(m, lr, msn, 0, 0) for (m, lr, msn) in params])
        check(lib().fitsne_batch_step(self.h, arr))

    def close(self):
        if self.h:
            lib().fitsne_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- learned β warm start

MLP_IN, MLP_HID = 96, 64
MLP_PARAMS = 2 * MLP_IN + MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1
MLP_TRAIN0 = 2 * MLP_IN


def mlp_unpack(p):
    """Split the FP32 parameter vector into named arrays (views)."""
    o = 0
    out = {}
    for name, shape in [("mu", (MLP_IN,)), ("sd", (MLP_IN,)), ("W1", (MLP_IN, MLP_HID)), ("b1", (MLP_HID,)),
                        ("W2", (MLP_HID, MLP_HID)), ("b2", (MLP_HID,)), ("w3", (MLP_HID,)), ("b3", (1,))]:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


class BetaModel:
    """MLP warm start of the β bisection (BF16 MFMA on the GPU)."""

    def __init__(self, K=90, seed=1):
        h = _vp()
        check(lib().fitsne_beta_model_create(C.byref(h), K, seed))
        self.h = h
        self.K = K

    def close(self):
        if self.h:
            lib().fitsne_beta_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fit(self, dist, beta, epochs=30, batch=1024, lr=1e-3, seed=7):
        dist = np.ascontiguousarray(dist, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        loss = _cd()
        check(lib().fitsne_beta_model_fit(self.h, dist.reshape(-1), beta, dist.shape[0], epochs, batch, lr, seed,
                                          C.byref(loss)))
        return loss.value

    def predict(self, dist):
        dist = np.ascontiguousarray(dist, dtype=np.float64)
        out = np.empty(dist.shape[0])
        check(lib().fitsne_beta_model_predict(self.h, dist.reshape(-1), dist.shape[0], out))
        return out

    def gradient(self, dist, beta):
        dist = np.ascontiguousarray(dist, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        g = np.empty(MLP_PARAMS - MLP_TRAIN0, dtype=np.float32)
        loss = _cd()
        check(lib().fitsne_beta_model_gradient(self.h, dist.reshape(-1), beta, dist.shape[0], g, C.byref(loss)))
        return g, loss.value

    def params(self):
        p = np.empty(MLP_PARAMS, dtype=np.float32)
        check(lib().fitsne_beta_model_params(self.h, p))
        return p

    def set_params(self, p):
        check(lib().fitsne_beta_model_set_params(self.h, np.ascontiguousarray(p, dtype=np.float32)))


# ---------------------------------------------------------------- resident tier

class Context:
    """Device-resident P, Y, momentum and gains (one optimisation iteration moves no data over PCIe)."""

    def __init__(self, no_dims=2, precision=FP32, n_interpolation_points=3, intervals_per_integer=1.0,
                 min_num_intervals=50, df=1.0, device=0):
        h = _vp()
        check(lib().fitsne_ctx_create(C.byref(h), device, precision, no_dims, n_interpolation_points,
                                      intervals_per_integer, min_num_intervals, df))
        self.h = h
        self.D = no_dims
        self.N = 0
        self.precision = precision

    def close(self):
        if self.h:
            lib().fitsne_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_P(self, row, col, val):
        N = row.size - 1
        val = np.ascontiguousarray(val, dtype=np.float64)
        if row.dtype == np.int64:
            check(lib().fitsne_set_P_i64(self.h, N, row, np.ascontiguousarray(col, dtype=np.int32), val))
        else:
            check(lib().fitsne_set_P(self.h, N, np.ascontiguousarray(row, dtype=np.uint32),
                                     np.ascontiguousarray(col, dtype=np.uint32), val))
        self.N = N

    def exact_affinities(self, X, perplexity=30.0, sigma=-1.0, perplexity_list=None, download=True):
        """Dense P of the exact mode built on the device (FP64_EXACT contexts).  Returns (P or None,
        sum_P before normalisation)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        N, D = X.shape
        plist = None if perplexity_list is None else np.ascontiguousarray(perplexity_list, dtype=np.float64)
        P = np.empty((N, N)) if download else None
        s = _cd()
        check(lib().fitsne_exact_affinities(self.h, X.reshape(-1), N, D, perplexity, sigma,
                                            0 if plist is None else plist.size, _ptr(plist), _ptr(P),
                                            C.byref(s)))
        self.N = N
        return P, s.value

    def set_P_dense(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        check(lib().fitsne_set_P_dense(self.h, P.shape[0], P.reshape(-1)))
        self.N = P.shape[0]

    def set_theta(self, theta):
        """Barnes-Hut opening parameter (FP64_BH contexts)."""
        check(lib().fitsne_set_theta(self.h, theta))

    def set_exaggeration(self, f):
        check(lib().fitsne_set_exaggeration(self.h, f))

    def set_reorder(self, enable):
        check(lib().fitsne_set_reorder(self.h, int(enable)))

    def set_graph(self, enable):
        """Replay steps as one captured graph with device-planned geometry (FP32 form)."""
        check(lib().fitsne_set_graph(self.h, int(enable)))

    def set_overlap(self, enable):
        """Attractive term and grid stages side by side on two streams, and in 1D the convolution's kernel
        line on a third (default), or all kernels in order on one stream.  Results are the same."""
        check(lib().fitsne_set_overlap(self.h, int(enable)))

    def autotune(self):
        """Time the spread's points per workgroup on the current Y and keep the fastest; returns it."""
        out = C.c_int(0)
        check(lib().fitsne_autotune(self.h, C.byref(out)))
        return out.value

    def tuned(self):
        """(points per spread workgroup, rows per attractive workgroup) now in use."""
        a, b = C.c_int(0), C.c_int(0)
        check(lib().fitsne_autotune_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_attractive_rows(self, rows):
        """Rows per attractive workgroup (0: automatic, -1: balanced blocks); results do not depend on it."""
        check(lib().fitsne_set_attractive_rows(self.h, int(rows)))

    def set_spread_points(self, points):
        """Points per spread workgroup (a multiple of 512 in [512, 8192]); results do not depend on it."""
        check(lib().fitsne_set_spread_points(self.h, int(points)))

    def reorder(self):
        check(lib().fitsne_reorder(self.h))

    def set_Y(self, Y):
        check(lib().fitsne_set_Y(self.h, np.ascontiguousarray(Y, dtype=np.float64).reshape(-1)))

    def get_Y(self):
        out = np.empty(self.N * self.D)
        check(lib().fitsne_get_Y(self.h, out))
        return out.reshape(self.N, self.D)

    def gradient(self, download=True):
        sq = _cd()
        dC = np.empty(self.N * self.D) if download else None
        check(lib().fitsne_gradient(self.h, _ptr(dC), C.byref(sq)))
        return (dC.reshape(self.N, self.D) if download else None), sq.value

    def step(self, momentum, learning_rate, max_step_norm=5.0, plain_descent=False, want_sumq=False):
        prm = StepParams(momentum, learning_rate, max_step_norm, int(plain_descent), 0)
        if want_sumq:
            sq = _cd()
            check(lib().fitsne_step(self.h, C.byref(prm), C.byref(sq)))
            return sq.value
        check(lib().fitsne_step(self.h, C.byref(prm), None))
        return None

    def kl(self):
        out = _cd()
        check(lib().fitsne_kl(self.h, C.byref(out)))
        return out.value

    def sync(self):
        check(lib().fitsne_sync(self.h))

    def set_timing(self, on):
        check(lib().fitsne_set_timing(self.h, int(on)))

    def kernel_stats(self, cap=64):
        names = C.create_string_buffer(48 * cap)
        counts = np.zeros(cap, dtype=np.int32)
        ms = np.zeros(cap)
        n = check(lib().fitsne_kernel_stats(self.h, cap, names, _ptr(counts), _ptr(ms)))
        raw = names.raw
        out = {}
        for i in range(n):
            nm = raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (int(counts[i]), float(ms[i]))
        return out

    def debug_read(self, name, count, dtype):
        out = np.zeros(count, dtype=dtype)
        n = check(lib().fitsne_debug_read(self.h, name.encode(), _ptr(out), out.nbytes))
        return out[: n // out.itemsize]

    def last_geometry(self):
        out = np.zeros(6)
        check(lib().fitsne_last_geometry(self.h, out))
        return dict(n_boxes=int(out[0]), n1d=int(out[1]), fft_len=int(out[2]), min=out[3], max=out[4],
                    nnz=int(out[5]))
