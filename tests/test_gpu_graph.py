"""Captured steps (device-planned geometry, one hipGraph per step) and batched embeddings on the GPU.

The captured step runs the same kernels on the same data as the host-planned step; only the launch
shapes (sized for the build's largest grid) and the place where the geometry is computed differ, and
the geometry rules are the same functions on host and device.  So after any number of steps the
embedding must be bitwise the host-planned one, in 2D and 1D, with and without the spatial reorder,
and each embedding of a batch bitwise its own single-context run.  A grid beyond the build's limits
is refused on the device and leaves the embedding unchanged.
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import synth

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[graph test] {msg}", flush=True)


def _ctx(P, Y, D, df=1.0, graph=False):
    c = fg.Context(no_dims=D, precision=fg.FP32, df=df)
    c.set_P(*P)
    c.set_Y(Y)
    c.N = Y.shape[0]
    if graph:
        c.set_graph(True)
    return c


def _schedule(k):
    return 0.5 if k < 6 else 0.8


@pytest.mark.parametrize("N,D,df", [(5000, 2, 1.0), (70000, 2, 1.0), (20000, 1, 0.5)])
def test_captured_steps_bitwise_equal_host_planned(N, D, df):
    P, Y, _ = synth.small_mixture_workload(N=N, D=D, span=40.0 if D == 2 else 200.0, seed=3)
    _note(f"N={N} D={D} df={df}")
    a = _ctx(P, Y, D, df)
    b = _ctx(P, Y, D, df, graph=True)
    lr = max(200.0, N / 12.0)
    for k in range(12):
        za = a.step(_schedule(k), lr, 5.0, want_sumq=(k % 5 == 4))
        zb = b.step(_schedule(k), lr, 5.0, want_sumq=(k % 5 == 4))
        assert za == zb
    ya, yb = a.get_Y(), b.get_Y()
    _note(f"max |dY| = {np.abs(ya - yb).max():.3e}")
    assert np.array_equal(ya, yb)
    assert a.kl() == b.kl()
    ga, gb = a.last_geometry(), b.last_geometry()
    assert ga["n_boxes"] == gb["n_boxes"] and ga["min"] == gb["min"] and ga["max"] == gb["max"]


def test_batch_matches_single_contexts():
    sizes = [2000, 3500, 7000, 12000, 26000, 41000]
    data = [synth.small_mixture_workload(N=n, D=2, span=30.0 + n / 1000.0, seed=10 + i) for i, n in enumerate(sizes)]
    _note(f"batch of {len(sizes)}: N = {sizes}")
    singles = [_ctx(P, Y, 2) for (P, Y, _) in data]
    members = [_ctx(P, Y, 2) for (P, Y, _) in data]
    batch = fg.Batch(members)
    for k in range(10):
        prm = [(_schedule(k), max(200.0, n / 12.0), 5.0) for n in sizes]
        batch.step(prm)
        for c, (m, lr, msn) in zip(singles, prm):
            c.step(m, lr, msn)
    for i, (s, m) in enumerate(zip(singles, members)):
        ys, ym = s.get_Y(), m.get_Y()
        assert np.array_equal(ys, ym), (sizes[i], np.abs(ys - ym).max())
    batch.close()


def test_grid_beyond_the_limit_is_refused_on_the_device():
    P, Y, _ = synth.small_mixture_workload(N=3000, D=2, span=40.0, seed=5)
    Y = Y.copy()
    Y[17] = [900.0, -900.0]  # span 1800 -> 1800 boxes per dimension, beyond 2048 / (2 * 3)
    c = _ctx(P, Y, 2, graph=True)
    before = c.get_Y()
    c.step(0.5, 200.0, 5.0)
    with pytest.raises(fg.FitsneError) as e:
        c.sync()
    assert e.value.code == fg.E_UNSUPPORTED
    assert np.array_equal(c.get_Y(), before)
    host = _ctx(P, Y, 2)
    with pytest.raises(fg.FitsneError):
        host.step(0.5, 200.0, 5.0)
