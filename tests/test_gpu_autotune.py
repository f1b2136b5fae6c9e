"""The autotuner's knobs, the spread's points per workgroup and the attractive term's rows per workgroup,
leave every result bitwise unchanged.

The spread's waves are the same aligned groups of 128 points for any workgroup size, and everything
after the in-wave sums is exact integer arithmetic, so the FP32 gradient, Z and a run of optimiser
steps (host-planned and captured) must be identical for 512, 1024, 2048 and 4096 points per
workgroup.  The autotuner must return one of those and leave the results as they were.
"""
import numpy as np
import pytest

import fitsne_gpu as fg
from datagen import blobs, random_csr

pytestmark = pytest.mark.gpu


def _run(P, Y, pts, D, graph=False, steps=4):
    c = fg.Context(no_dims=D, precision=fg.FP32)
    c.set_P(*P)
    c.set_Y(Y)
    if pts is None:
        chosen = c.autotune()
    else:
        c.set_spread_points(pts)
        chosen = pts
    g, z = c.gradient()
    if graph:
        c.set_graph(True)
    for _ in range(steps):
        c.step(momentum=0.5, learning_rate=200.0)
    Yout = c.get_Y()
    c.close()
    return chosen, g, z, Yout


@pytest.mark.parametrize("D,N", [(2, 30_000), (1, 20_000)])
def test_spread_workgroup_size_does_not_change_results(D, N):
    Y = blobs(N, D, 6, seed=31 + D, span=40.0)
    P = random_csr(N, 15, seed=5)
    print(f"[autotune test] D={D} N={N}", flush=True)
    _, g0, z0, y0 = _run(P, Y, 1024, D)
    for pts in (512, 2048, 4096):
        _, g, z, y = _run(P, Y, pts, D)
        np.testing.assert_array_equal(g, g0)
        assert z == z0
        np.testing.assert_array_equal(y, y0)
    _, g, z, y = _run(P, Y, 512, D, graph=True)
    _, gr, zr, yr = _run(P, Y, 1024, D, graph=True)
    np.testing.assert_array_equal(y, yr)
    chosen, g, z, y = _run(P, Y, None, D)
    assert chosen in (512, 1024, 2048, 4096)
    np.testing.assert_array_equal(g, g0)
    np.testing.assert_array_equal(y, y0)


@pytest.mark.parametrize("D,N", [(2, 30_000), (1, 20_000)])
def test_attractive_row_block_does_not_change_results(D, N):
    """The attractive term's rows per workgroup, fixed or balanced by the cost model (-1), change only
    which workgroup sums which rows: with one fixed-point scale for all rows the forces, and so every
    step, are bitwise the same."""
    Y = blobs(N, D, 6, seed=41 + D, span=40.0)
    P = random_csr(N, 15, seed=6)
    out = []
    cand = (0, 64, 1024, 4096, -1) + ((8192,) if D == 1 else ())  # 1D blocks reach 8192 rows
    for rows in cand:
        c = fg.Context(no_dims=D, precision=fg.FP32)
        c.set_P(*P)
        c.set_Y(Y)
        c.set_attractive_rows(rows)
        g, z = c.gradient()
        for _ in range(3):
            c.step(momentum=0.5, learning_rate=200.0)
        out.append((g, z, c.get_Y(), c.tuned()[1]))
        c.close()
    for g, z, y, r in out[1:]:
        np.testing.assert_array_equal(g, out[0][0])
        assert z == out[0][1]
        np.testing.assert_array_equal(y, out[0][2])
    assert [o[3] for o in out[1:]] == list(cand[1:])


def test_attractive_rows_are_validated():
    """Row blocks are limited by the kernel's LDS: 4096 rows in 2D, 8192 in 1D, multiples of 64."""
    for D, top in ((2, 4096), (1, 8192)):
        c = fg.Context(no_dims=D, precision=fg.FP32)
        c.set_attractive_rows(top)
        for bad in (-2, 32, 100, top + 64):
            with pytest.raises(fg.FitsneError):
                c.set_attractive_rows(bad)
        c.close()


def test_spread_points_are_validated():
    c = fg.Context(no_dims=2, precision=fg.FP32)
    for bad in (0, 100, 256, 300, 768, 16384):
        with pytest.raises(fg.FitsneError):
            c.set_spread_points(bad)
    with pytest.raises(fg.FitsneError):
        c.autotune()  # no P and Y yet
    c.close()


def test_balanced_row_blocks_at_scale():
    """Blocks cut by the cost model (-1) at a size where they differ from the fixed ones (up to 4096 rows
    each, one per CU): the gradient and Z are bitwise those of the automatic row blocks."""
    N = 400_000
    Y = blobs(N, 2, 8, seed=51, span=60.0)
    P = random_csr(N, 20, seed=7)
    print(f"[autotune test] balanced blocks N={N}", flush=True)
    res = []
    for rows in (0, -1):
        c = fg.Context(no_dims=2, precision=fg.FP32)
        c.set_P(*P)
        c.set_Y(Y)
        c.set_attractive_rows(rows)
        g, z = c.gradient()
        res.append((g, z, c.tuned()[1]))
        c.close()
    np.testing.assert_array_equal(res[1][0], res[0][0])
    assert res[1][1] == res[0][1]
    assert res[1][2] == -1
