"""GPU parity of the HIP path against the oracle on identical seeded inputs (through the C-ABI).

Tolerances: FP64 plain path ||dC - dC_oracle|| / ||dC_oracle|| <= 1e-9 (different FFT length and
summation order only); FP32 form <= 1e-4 (the target bound; measured values are printed).
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import orc
from accuracy import gradient_errors
from datagen import blobs, random_csr

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def say(*a):
    print(*a, flush=True)


def test_interpolate_matches_oracle():
    say("[gpu] interpolate p=3,4")
    rng = np.random.default_rng(0)
    y = rng.uniform(-0.05, 1.05, 1000)
    for p in (3, 4):
        sp = np.empty(p)
        sp[0] = 0.5 / p
        for i in range(1, p):
            sp[i] = sp[i - 1] + 1.0 / p
        got = fg.interpolate(p, y, sp)
        want = orc.interpolate(p, y, sp)
        assert np.max(np.abs(got - want)) <= 1e-14


def test_distances2similarities_matches_oracle():
    say("[gpu] distances2similarities n=500 K=90")
    rng = np.random.default_rng(1)
    n, K = 500, 90
    Dist = np.sort(rng.gamma(2.0, 2.0, size=(n, K)), axis=1)
    P, beta, iters = fg.distances2similarities(Dist, 30.0, squared=False)
    import ctypes as C
    for i in range(0, n, 7):
        Pw = np.empty(K)
        it = C.c_int()
        H = C.c_double()
        bw = orc.oracle().orc_d2s_bisect(Dist[i].copy(), Pw, K, 30.0, 1.0, 0, C.byref(it), C.byref(H))
        assert abs(beta[i] - bw) <= 1e-12 * abs(bw)
        assert iters[i] == it.value
        assert np.max(np.abs(P[i] - Pw)) <= 1e-13


@pytest.mark.parametrize("D,df", [(2, 1.0), (2, 0.5), (1, 1.0), (1, 0.5)])
def test_mirror_gradient_fp64_matches_oracle(D, df):
    say(f"[gpu] mirror fp64 gradient D={D} df={df} N=3000")
    N = 3000
    P = random_csr(N, 12, seed=10 + D)
    Y = blobs(N, D, 7, seed=3, span=40.0, sd=2.0)
    want, zq = orc.gradient(P, Y, df=df)
    got, sq = fg.compute_fft_gradient(P, Y, df=df)
    say(f"    rel={rel(got, want):.3e} Zrel={abs(sq - zq) / zq:.3e}")
    assert abs(sq - zq) / abs(zq) <= 1e-9
    assert rel(got, want) <= 1e-9


@pytest.mark.parametrize("D,df", [(2, 1.0), (2, 0.5), (1, 1.0), (1, 0.5)])
def test_fp32_gradient_matches_oracle(D, df):
    say(f"[gpu] fp32 gradient D={D} df={df} N=5000")
    N = 5000
    P = random_csr(N, 12, seed=20 + D)
    Y = blobs(N, D, 9, seed=5, span=60.0, sd=2.5)
    # the FP32 context works on the FP32-rounded embedding; compare on the same inputs
    Y32 = Y.astype(np.float32).astype(np.float64)
    want, zq = orc.gradient(P, Y32, df=df)
    ctx = fg.Context(no_dims=D, precision=fg.FP32, df=df)
    ctx.set_P(*P)
    ctx.set_Y(Y)
    got, sq = ctx.gradient()
    ctx.close()
    r = rel(got, want)
    e = gradient_errors(got, want)
    say(f"    rel={r:.3e} Zrel={abs(sq - zq) / zq:.3e} per point: max {e['max_rel']:.3e} p50 {e['p50']:.3e} "
        f"p99 {e['p99']:.3e}")
    assert abs(sq - zq) / abs(zq) <= 1e-5
    assert r <= 1e-4
    # per point: the largest error against the largest gradient, and the 99th percentile of each point's
    # error against its own gradient (plus 1e-3 of the rms gradient)
    assert e["max_rel"] <= 1e-4
    assert e["p99"] <= 1e-3


def test_n_body_fft_2d_mirror_matches_oracle():
    say("[gpu] n_body_fft_2d mirror (precompute_2d + n_body_fft_2d)")
    rng = np.random.default_rng(4)
    N, nb, p = 2000, 36, 3
    Y = rng.uniform(-12, 15, size=(N, 2))
    mn, mx = orc.minmax_quirk(Y)
    xs, ys = Y[:, 0].copy(), Y[:, 1].copy()
    ch = np.stack([np.ones(N), xs, ys, xs * xs + ys * ys], axis=1)
    for kernel, df in ((orc.SQUARED_CAUCHY, 1.0), (orc.GENERAL, 0.7)):
        go = orc.precompute_2d(mx, mn, nb, p, kernel, df)
        want = orc.n_body_fft_2d(xs, ys, ch, nb, p, go)
        gg = fg.precompute_2d(mx, mn, nb, p, kernel, df)
        for k in ("lo", "hi", "sp", "yt", "xt"):
            assert np.array_equal(gg[k], go[k]), k
        got = fg.n_body_fft_2d(xs, ys, ch, nb, p, gg, potential=np.ones((N, 4)))
        err = np.abs(got - 1.0 - want).max(axis=0) / np.abs(want).max(axis=0)
        say(f"    kernel={kernel} max rel per term {err}")
        assert np.all(err <= 1e-11)


def test_n_body_fft_1d_mirror_matches_oracle():
    say("[gpu] n_body_fft_1d mirror")
    rng = np.random.default_rng(5)
    N, nb, p = 3000, 120, 3
    Y = rng.uniform(-40, 70, size=N)
    ch = np.stack([np.ones(N), Y, Y * Y], axis=1)
    for kernel, df in ((orc.SQUARED_CAUCHY, 1.0), (orc.SQUARED_GENERAL, 0.5)):
        go = orc.precompute_1d(Y.min(), Y.max(), nb, p, kernel, df)
        want = orc.n_body_fft_1d(Y, ch, nb, p, go)
        gg = fg.precompute_1d(Y.min(), Y.max(), nb, p, kernel, df)
        for k in ("lo", "hi", "sp", "yt"):
            assert np.array_equal(gg[k], go[k]), k
        got = fg.n_body_fft_1d(Y, ch, nb, p, gg)
        err = np.abs(got - want).max(axis=0) / np.abs(want).max(axis=0)
        say(f"    kernel={kernel} max rel per term {err}")
        assert np.all(err <= 1e-11)


def test_kl_matches_oracle():
    say("[gpu] evaluate_error_fft")
    N = 4000
    P = random_csr(N, 10, seed=7)
    Y = blobs(N, 2, 6, seed=8, span=30.0)
    for df in (1.0, 0.6):
        want = orc.kl(P, Y, 5432.1, df=df)
        got = fg.evaluate_error_fft(P, Y, 5432.1, df=df)
        assert abs(got - want) <= 1e-12 * abs(want)


def test_fp64_steps_follow_the_oracle_update():
    say("[gpu] fp64 plain: 3 optimiser steps vs oracle")
    N = 2000
    P = random_csr(N, 10, seed=9)
    Y = blobs(N, 2, 5, seed=1, span=25.0)
    ctx = fg.Context(no_dims=2, precision=fg.FP64_PLAIN)
    ctx.set_P(*P)
    ctx.set_exaggeration(12.0)
    ctx.set_Y(Y)
    Yo = Y.copy().reshape(-1)
    uY = np.zeros_like(Yo)
    gains = np.ones_like(Yo)
    row, col, val = P
    valx = val * 12.0
    for it in range(3):
        ctx.step(0.5, 200.0, 5.0)
        dC, _ = orc.gradient((row, col, valx), Yo.reshape(N, 2))
        orc.oracle().orc_update(Yo, uY, gains, dC.reshape(-1).copy(), N, 2, 0.5, 200.0, 5.0, 0)
        orc.oracle().orc_zero_mean(Yo, N, 2)
    got = ctx.get_Y()
    ctx.close()
    r = rel(got.reshape(-1), Yo)
    say(f"    Y rel after 3 steps {r:.3e}")
    assert r <= 1e-9


def test_minmax_quirk_when_first_coordinate_is_the_minimum():
    say("[gpu] 2D min/max else-if rule with xs[0] the strict minimum")
    N = 1500
    P = random_csr(N, 8, seed=11)
    Y = blobs(N, 2, 4, seed=12, span=20.0)
    Y[0, 0] = Y.min() - 3.0
    for prec, tol in ((fg.FP64_PLAIN, 1e-9), (fg.FP32, 1e-4)):
        Yin = Y if prec == fg.FP64_PLAIN else Y.astype(np.float32).astype(np.float64)
        want, _ = orc.gradient(P, Yin)
        ctx = fg.Context(no_dims=2, precision=prec)
        ctx.set_P(*P)
        ctx.set_Y(Y)
        got, _ = ctx.gradient()
        geo = ctx.last_geometry()
        ctx.close()
        mn, mx = orc.minmax_quirk(Yin)
        assert geo["min"] == mn and geo["max"] == mx
        assert rel(got, want) <= tol


def test_fp32_bitwise_repeatable_across_contexts():
    """Regression for device buffers reused between contexts: counters and grids must be initialised on
    the context's own stream, so two fresh contexts (and two calls on one) give identical bits."""
    say("[gpu] fp32 repeatability across contexts")
    N = 4000
    P = random_csr(N, 10, seed=31)
    Y = blobs(N, 2, 6, seed=32, span=50.0)
    outs = []
    for _ in range(3):
        ctx = fg.Context(no_dims=2, precision=fg.FP32)
        ctx.set_P(*P)
        ctx.set_Y(Y)
        g1, z1 = ctx.gradient()
        g2, z2 = ctx.gradient()
        ctx.close()
        assert np.array_equal(g1, g2) and z1 == z2
        outs.append((g1, z1))
    for g, z in outs[1:]:
        assert np.array_equal(g, outs[0][0]) and z == outs[0][1]


def test_fp32_spatial_reorder_is_invisible():
    """The FP32 context keeps points in spatial order internally (N >= 65536) and then gathers attractive
    neighbours through a column-blocked copy of P.  Results come back in the caller's order, agree with
    the unreordered path to FP32 summation order (1e-6), are bitwise repeatable, and follow the
    reference's `else if` min/max rule for the caller's first point."""
    say("[gpu] fp32 reorder N=70000")
    from synth import small_mixture_workload
    P, Y, _ = small_mixture_workload(N=70000, D=2, k=12, per_point=20, span=60.0, seed=3)
    Y[0, 0] = Y.min() - 1.0  # the caller's xs[0] is the strict minimum: `else if` rule in play
    outs = []
    for reorder in (True, True, False):
        ctx = fg.Context(no_dims=2, precision=fg.FP32)
        ctx.set_reorder(reorder)
        ctx.set_P(*P)
        ctx.set_Y(Y)
        g, z = ctx.gradient()
        geo = ctx.last_geometry()
        back = ctx.get_Y()
        ctx.close()
        assert np.array_equal(back, Y.astype(np.float32).astype(np.float64))
        outs.append((g, z, geo))
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]
    r01 = rel(outs[0][0], outs[2][0])
    say(f"    reordered vs caller order rel {r01:.3e}")
    assert r01 <= 1e-6 and abs(outs[0][1] - outs[2][1]) <= 1e-6 * abs(outs[2][1])
    mn, mx = orc.minmax_quirk(Y.astype(np.float32).astype(np.float64))
    assert outs[0][2]["min"] == mn and outs[0][2]["max"] == mx
    Pu = (P[0].astype(np.uint32), P[1].astype(np.uint32), P[2])
    want, zq = orc.gradient(Pu, Y.astype(np.float32).astype(np.float64), nthreads=8)
    r = rel(outs[0][0], want)
    say(f"    rel vs oracle {r:.3e}")
    assert r <= 1e-4


def test_fp32_steps_track_the_oracle():
    """Ten full iterations (gradient, gains/momentum, clip, zeroMean) of the FP32 form against the oracle's
    FP64 iterations from the same start: the trajectories stay within FP32 rounding."""
    say("[gpu] fp32 10 steps vs oracle")
    N = 3000
    P = random_csr(N, 10, seed=41)
    Y = blobs(N, 2, 5, seed=42, span=30.0)
    Y = Y.astype(np.float32).astype(np.float64)
    ctx = fg.Context(no_dims=2, precision=fg.FP32)
    ctx.set_P(*P)
    ctx.set_exaggeration(12.0)
    ctx.set_Y(Y)
    Yo = Y.copy().reshape(-1)
    uY = np.zeros_like(Yo)
    gains = np.ones_like(Yo)
    row, col, val = P
    for it in range(10):
        ctx.step(0.5, 200.0, 5.0)
        dC, _ = orc.gradient((row, col, val * 12.0), Yo.reshape(N, 2))
        orc.oracle().orc_update(Yo, uY, gains, dC.reshape(-1).copy(), N, 2, 0.5, 200.0, 5.0, 0)
        orc.oracle().orc_zero_mean(Yo, N, 2)
    got = ctx.get_Y()
    ctx.close()
    r = rel(got.reshape(-1), Yo)
    say(f"    Y rel after 10 steps {r:.3e}")
    assert r <= 1e-4
