"""The benchmark configurations at full size on the GPU, where the oracle is out of reach.

At C2 (N = 1,000,000, 2D, df = 1) and C3 (N = 2,000,000, 1D, df = 0.5) the FP32 form must agree with
the FP64 plain path on the same Y and P (gradient relative error <= 1e-4, Z <= 1e-5; the plain path
agrees with the oracle to ~1e-12 at the sizes the oracle can run), and it must be bitwise repeatable.
The geometry is checked against the reference's rules: C2 -> 175 boxes per dimension (span 160, rounding
table of tsne.cpp:1071-1077), C3 -> 600 boxes (1D: span / intervals_per_integer, no rounding).
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import synth

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[scale test] {msg}", flush=True)


def _grad(P, Y, precision, df):
    c = fg.Context(no_dims=Y.shape[1], precision=precision, df=df)
    c.set_P(*P)
    c.set_Y(Y)
    g, z = c.gradient()
    geo = c.last_geometry()
    c.close()
    return g, z, geo


@pytest.mark.parametrize("name", ["C2", "C3"])
def test_full_size_fp32_matches_fp64_plain(name):
    if name == "C2":
        P, Y, _ = synth.c2_workload()
        df, nb = 1.0, 175
    else:
        P, Y, _ = synth.c3_workload()
        df, nb = 0.5, None
    _note(f"{name}: N={Y.shape[0]} D={Y.shape[1]} nnz={P[2].size} df={df}")
    Y32 = Y.astype(np.float32).astype(np.float64)  # both paths on the same (FP32-representable) Y
    g32, z32, geo = _grad(P, Y32, fg.FP32, df)
    g32b, z32b, _ = _grad(P, Y32, fg.FP32, df)
    g64, z64, geo64 = _grad(P, Y32, fg.FP64_PLAIN, df)
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    _note(f"{name}: boxes {geo['n_boxes']} n1d {geo['n1d']} fft {geo['fft_len']}; grad rel {rel:.3e}, "
          f"Z rel {abs(z32 - z64) / abs(z64):.3e}")
    assert geo["n_boxes"] == geo64["n_boxes"]
    if nb is not None:
        assert geo["n_boxes"] == nb
    else:
        span = geo["max"] - geo["min"]
        assert geo["n_boxes"] == max(50, int(span))
    assert rel <= 1e-4
    assert abs(z32 - z64) / abs(z64) <= 1e-5
    assert np.array_equal(g32, g32b) and z32 == z32b
