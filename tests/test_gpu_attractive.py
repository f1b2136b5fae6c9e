"""The FP32 form's windowed attractive kernel, isolated and checked row by row against the oracle.

In the FP32 form the exaggeration multiplies the attractive forces only, so with the same Y,
dC(exaggeration 2) - dC(exaggeration 1) is the attractive term (tsne.cpp:1121-1137) as the kernel
computed it, up to the rounding of the two gradients.  Every row is checked against
orc_attractive with a bound of 1e-5 of the row's sum of absolute terms: a dropped or duplicated entry
moves its row by at least a 300th of that sum.

The affinities cover the layout's cases: rows longer than one 64-entry chunk (up to 300 entries),
rows spread over many column windows of 8192 points, empty rows and rows of one entry.  With the
spatial reorder most entries sit in row-pure quads; without it (columns scattered over the windows)
most sit in the mixed sections.
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import orc
from datagen import blobs

pytestmark = pytest.mark.gpu


def _affinities(N, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, N)
    long_rows = rng.random(N) < 0.05
    lens[long_rows] = rng.integers(65, 301, int(long_rows.sum()))
    lens[rng.random(N) < 0.05] = 0
    lens[rng.random(N) < 0.05] = 1
    cols = []
    for n in range(N):
        k = int(lens[n])
        near = n + rng.integers(-3000, 3001, k)  # points are sorted by x: near in the plane
        far = rng.integers(0, N, k)
        c = np.where(rng.random(k) < 0.8, near, far) % N
        c = np.unique(c[c != n])
        cols.append(c)
    row = np.zeros(N + 1, np.uint32)
    row[1:] = np.cumsum([c.size for c in cols])
    col = np.concatenate(cols).astype(np.uint32)
    val = rng.uniform(0.5, 1.0, col.size)
    return row, col, val / val.sum()


def _abs_terms(P, Y):
    row, col, val = P
    r = np.repeat(np.arange(Y.shape[0]), np.diff(row.astype(np.int64)))
    d = Y[r] - Y[col.astype(np.int64)]
    q = 1.0 / (1.0 + (d * d).sum(1))
    return np.bincount(r, weights=val * q * np.sqrt((d * d).sum(1)), minlength=Y.shape[0])


@pytest.mark.parametrize("reorder", [True, False])
def test_windowed_attractive_term_row_by_row(reorder):
    N = 120_000
    Y = blobs(N, 2, 12, seed=21, span=60.0)
    Y = Y[np.argsort(Y[:, 0], kind="stable")]
    P = _affinities(N, 22)
    print(f"[attractive test] N={N} nnz={int(P[0][-1])} reorder={reorder}", flush=True)
    c = fg.Context(no_dims=2, precision=fg.FP32)
    c.set_reorder(reorder)
    c.set_P(*P)
    c.set_Y(Y)
    g1, _ = c.gradient()
    c.set_exaggeration(2.0)
    g2, _ = c.gradient()
    c.close()
    Y32 = Y.astype(np.float32).astype(np.float64)
    want = orc.attractive(P, Y32)
    got = g2 - g1
    scale = _abs_terms(P, Y32) + np.abs(g1).sum(1) + np.abs(g2).sum(1) + 1e-30
    err = np.abs(got - want).max(1) / scale
    assert err.max() <= 1e-5, f"worst row {int(err.argmax())}: {err.max():.3e}"
    # rows without entries: both gradients are the same repulsion, bit for bit
    empty = np.diff(P[0].astype(np.int64)) == 0
    assert empty.any() and (got[empty] == 0).all()
