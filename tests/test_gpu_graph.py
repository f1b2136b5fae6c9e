"""Captured steps (device-planned geometry, one hipGraph per step) and batched embeddings on the GPU.

The captured step runs the same kernels on the same data as the host-planned step; only the launch
shapes (sized for the build's largest grid) and the place where the geometry is computed differ, and
the geometry rules are the same functions on host and device (in 1D the host-planned step transforms the
kernel line in a workgroup of its own, the captured step inside the convolution, with the same
per-element arithmetic).  So after any number of steps the
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


def _ctx(P, Y, D, df=1.0, graph=False, reorder=True):
    c = fg.Context(no_dims=D, precision=fg.FP32, df=df)
    c.set_P(*P)
    c.set_reorder(reorder)
    c.set_Y(Y)
    c.N = Y.shape[0]
    if graph:
        c.set_graph(True)
    return c


def _schedule(k):
    return 0.5 if k < 6 else 0.8


@pytest.mark.parametrize("N,D,df,reorder", [(5000, 2, 1.0, False), (5000, 2, 1.0, True), (70000, 2, 1.0, True),
                                             (20000, 1, 0.5, True), (20000, 1, 0.5, False)])
def test_captured_steps_bitwise_equal_host_planned(N, D, df, reorder):
    P, Y, _ = synth.small_mixture_workload(N=N, D=D, span=40.0 if D == 2 else 200.0, seed=3)
    _note(f"N={N} D={D} df={df} reorder={reorder}")
    a = _ctx(P, Y, D, df, reorder=reorder)
    b = _ctx(P, Y, D, df, graph=True, reorder=reorder)
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


def test_host_gradient_after_captured_steps_recomputes_kernel_spectra():
    """Host-planned and captured steps share the kernel-spectrum buffers: a host gradient at Y0, captured
    steps at other geometries, then Y0 again must give the gradient of a fresh context (the host step's
    record of which spectra it holds does not survive a captured step, a reset Y or a new capture)."""
    P, Y0, _ = synth.small_mixture_workload(N=6000, D=2, span=40.0, seed=8)
    _note("host gradient, captured steps, set_Y(Y0), host gradient")
    c = _ctx(P, Y0, 2)
    g0, z0 = c.gradient()
    c.set_graph(True)
    for k in range(4):
        c.step(0.5, 500.0, 5.0)  # large steps: the grid bounds move
    c.sync()
    assert c.last_geometry()["min"] != fg_min(Y0)
    c.set_graph(False)
    c.set_Y(Y0)
    g1, z1 = c.gradient()
    fresh = _ctx(P, Y0, 2)
    g2, z2 = fresh.gradient()
    assert np.array_equal(g1, g2) and z1 == z2
    assert np.array_equal(g0, g2) and z0 == z2


def fg_min(Y):
    return float(Y.min())


@pytest.mark.parametrize("D", [2, 1])
def test_update_steps_interleaved_with_other_calls_bitwise_equal(D):
    """Host-planned update steps leave zeroMean pending for the next step's readers of Y; every other
    entry point applies it first.  Steps interleaved with those calls (read Y, KL, gradient, captured
    steps, the autotuner, other workgroup sizes, Z requests) must give the uninterrupted steps' embedding
    bit for bit: none of the calls changes a result, only when the centring reaches memory."""
    N = 30000 if D == 2 else 40000
    P, Y, _ = synth.small_mixture_workload(N=N, D=D, span=40.0 if D == 2 else 200.0, seed=21)
    _note(f"interleaved calls, N={N} D={D}")
    a = _ctx(P, Y, D)
    b = _ctx(P, Y, D)
    lr = max(200.0, N / 12.0)
    for k in range(14):
        a.step(_schedule(k), lr, 5.0)
        if k == 3:
            b.set_graph(True)
        if k == 6:
            b.set_graph(False)
        zb = b.step(_schedule(k), lr, 5.0, want_sumq=(k % 4 == 1))
        if zb is not None:
            assert np.isfinite(zb) and zb > 0
        if k == 1:
            b.get_Y()
        elif k == 2:
            b.kl()
        elif k == 7:
            b.gradient()
        elif k == 8:
            b.autotune()
        elif k == 9:
            b.set_attractive_rows(1024)
        elif k == 10:
            b.set_spread_points(1024)
        elif k == 11:
            b.set_attractive_rows(-1)
    ya, yb = a.get_Y(), b.get_Y()
    _note(f"max |dY| = {np.abs(ya - yb).max():.3e}")
    assert np.array_equal(ya, yb)
    assert a.kl() == b.kl()
    a.close()
    b.close()


@pytest.mark.parametrize("D,df", [(2, 1.0), (1, 1.0), (1, 0.5)])
def test_side_streams_leave_results_bitwise_unchanged(D, df):
    """A host-planned step runs the attractive term and the grid stages side by side on two streams (in
    1D the attractive term stays on the context's stream and the grid stages move) and, in 1D, the
    convolution's kernel line on a third; with overlap off every kernel runs in turn on the context's
    stream.  Both
    must give the same embedding, Z and gradient bit for bit, also across a reset to a three times wider
    Y (a longer FFT line: the kernel-line and spectrum buffers grow) and back."""
    N = 20000
    P, Y, _ = synth.small_mixture_workload(N=N, D=D, span=40.0 if D == 2 else 200.0, seed=33)
    _note(f"overlap on / off, N={N} D={D} df={df}")
    a = _ctx(P, Y, D, df)
    b = _ctx(P, Y, D, df)
    b.set_overlap(False)
    lr = max(200.0, N / 12.0)
    lens = []
    for phase, Yp in enumerate([Y, 3.0 * Y, Y]):
        if phase:
            a.set_Y(Yp)
            b.set_Y(Yp)
        for k in range(5):
            za = a.step(_schedule(k), lr, 5.0, want_sumq=(k == 4))
            zb = b.step(_schedule(k), lr, 5.0, want_sumq=(k == 4))
            assert za == zb
        lens.append(a.last_geometry()["fft_len"])
        assert np.array_equal(a.get_Y(), b.get_Y()), phase
    _note(f"FFT lengths per phase: {lens}")
    assert lens[1] > lens[0]
    ga, za = a.gradient()
    gb, zb = b.gradient()
    assert np.array_equal(ga, gb) and za == zb
    a.close()
    b.close()


def test_fused_batch_follows_member_changes_and_mixed_batches_fall_back():
    """A member whose Y is reset between batched steps (new layout: the fused batch is rebuilt) still
    matches its single context bitwise, and so does a host-planned gradient taken from a member between
    batched steps.  A batch mixing 1D and 2D problems runs each member's own captured step instead."""
    sizes = [3000, 9000, 5000]
    data = [synth.small_mixture_workload(N=n, D=2, span=30.0, seed=30 + i) for i, n in enumerate(sizes)]
    singles = [_ctx(P, Y, 2) for (P, Y, _) in data]
    members = [_ctx(P, Y, 2) for (P, Y, _) in data]
    batch = fg.Batch(members)
    prm = [(0.5, 200.0, 5.0)] * len(sizes)
    for k in range(3):
        batch.step(prm)
        for c in singles:
            c.step(0.5, 200.0, 5.0)
    g_m, z_m = members[1].gradient()
    g_s, z_s = singles[1].gradient()
    assert np.array_equal(g_m, g_s) and z_m == z_s
    Ynew = data[2][1] * 1.5
    members[2].set_Y(Ynew)
    singles[2].set_Y(Ynew)
    for k in range(3):
        batch.step(prm)
        for c in singles:
            c.step(0.5, 200.0, 5.0)
    for s, m in zip(singles, members):
        assert np.array_equal(s.get_Y(), m.get_Y())
    batch.close()
    P1, Y1, _ = synth.small_mixture_workload(N=4000, D=1, span=100.0, seed=40)
    mixed = [_ctx(data[0][0], data[0][1], 2), _ctx(P1, Y1, 1, df=0.5)]
    alone = [_ctx(data[0][0], data[0][1], 2), _ctx(P1, Y1, 1, df=0.5)]
    mb = fg.Batch(mixed)
    for k in range(4):
        mb.step(prm[:2])
        for c in alone:
            c.step(0.5, 200.0, 5.0)
    for s, m in zip(alone, mixed):
        assert np.array_equal(s.get_Y(), m.get_Y())
    mb.close()
