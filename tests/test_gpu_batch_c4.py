"""The batched-embeddings workload at full size: 64 independent 2D problems with N log-uniform in
[2,000, 200,000] (synth.c4_problems, 2.6M points in all), stepped together by fitsne_batch_step.

Every member must be bitwise the embedding its own single context reaches with host-planned steps, after
steps across the momentum switch; the two smallest members must track the oracle's FP64 iterations
(gradient, gains/momentum, clip, zeroMean: tsne.cpp:440-531) from the same start within the FP32 form's
tolerance (relative error of Y <= 1e-4).
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import orc
import synth

pytestmark = pytest.mark.gpu

STEPS = 8


def _note(msg):
    print(f"[c4 test] {msg}", flush=True)


def _schedule(k):
    return 0.5 if k < 4 else 0.8


def _ctx(P, Y):
    c = fg.Context(no_dims=2, precision=fg.FP32)
    c.set_P(*P)
    c.set_Y(Y)
    c.N = Y.shape[0]
    return c


def test_c4_batch_matches_single_contexts_and_the_oracle():
    probs = synth.c4_problems()
    sizes = [Y.shape[0] for (_, Y) in probs]
    _note(f"{len(probs)} problems, N in [{min(sizes)}, {max(sizes)}], total {sum(sizes)}")
    lrs = [max(200.0, n / 12.0) for n in sizes]
    members = [_ctx(P, Y) for (P, Y) in probs]
    batch = fg.Batch(members)
    for k in range(STEPS):
        batch.step([(_schedule(k), lr, 5.0) for lr in lrs])
    batched = [m.get_Y() for m in members]
    batch.close()
    for m in members:
        m.close()
    _note("batched steps done; single contexts one at a time")
    for i, (P, Y) in enumerate(probs):
        c = _ctx(P, Y)
        for k in range(STEPS):
            c.step(_schedule(k), lrs[i], 5.0)
        ys = c.get_Y()
        c.close()
        assert np.array_equal(ys, batched[i]), (i, sizes[i], float(np.abs(ys - batched[i]).max()))
    for i in np.argsort(sizes)[:2]:
        (row, col, val), Y = probs[i]
        N = sizes[i]
        P32 = (row.astype(np.uint32), col.astype(np.uint32), val)
        Yo = Y.astype(np.float32).astype(np.float64).reshape(-1).copy()
        uY = np.zeros_like(Yo)
        gains = np.ones_like(Yo)
        for k in range(STEPS):
            dC, _ = orc.gradient(P32, Yo.reshape(N, 2))
            orc.oracle().orc_update(Yo, uY, gains, dC.reshape(-1).copy(), N, 2, _schedule(k), lrs[i], 5.0, 0)
            orc.oracle().orc_zero_mean(Yo, N, 2)
        r = float(np.linalg.norm(batched[i].reshape(-1) - Yo) / np.linalg.norm(Yo))
        _note(f"member {i} (N={N}) vs oracle after {STEPS} steps: Y rel {r:.3e}")
        assert r <= 1e-4
