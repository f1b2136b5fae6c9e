"""Pins the oracle's restatement of TSNE::run's optimisation loop (tsne.cpp:389-577) against the reference's
own TSNE::run, compiled from its sources (oracle/_ref/libfitsne_ref_run.so).

The reference's loop can run here through its Barnes-Hut (nbody_algorithm = 1) and exact (theta = 0)
branches, which do not reach FFTW.  The oracle keeps ONE loop (run_loop in oracle/orc_tsne.c) for its three
gradient sources, so what these runs pin -- the early exaggeration and its switch-off after the update of
stop_lying_iter, the late exaggeration, the momentum switch, no_momentum_during_exag's plain descent, the
step clip, zeroMean, which iterations record a KL and the exaggeration correction of that KL -- is the
same code that orc_run_fft runs around the FFT gradient.

Each case compares the final embedding and every recorded KL, bit for bit.  The run lengths are chosen so
that the last iteration (which always records a KL) falls before, on and after each switch.  A sensitivity
check shows that the comparison catches a switch moved by one iteration.
"""
import numpy as np
import pytest

import orc
from datagen import blobs, random_csr

pytestmark = pytest.mark.skipif(not orc.have_ref_run(), reason="oracle/_ref not built (reference sources absent)")

BASE = dict(perplexity=5.0, theta=0.5, stop_lying_iter=5, mom_switch_iter=8, momentum=0.5, final_momentum=0.8,
            learning_rate=150.0, early_exag_coeff=12.0, no_momentum_during_exag=0, start_late_exag_iter=11,
            late_exag_coeff=3.0, df=1.0, max_step_norm=5.0)


def reference_run(tmp, Y0, P=None, Pd=None, **kw):
    s = dict(BASE)
    s.update(kw)
    return orc.ref_tsne_run(tmp, Y0, s, P=P, Pd=Pd)


def oracle_kw(**kw):
    s = dict(BASE)
    s.update(kw)
    return {k: s[k] for k in ("stop_lying_iter", "mom_switch_iter", "momentum", "final_momentum", "learning_rate",
                              "early_exag_coeff", "no_momentum_during_exag", "start_late_exag_iter",
                              "late_exag_coeff", "df", "max_step_norm", "max_iter")}


def bh_case(N, nd, seed):
    P = random_csr(N, 6, seed=seed)
    Y0 = blobs(N, nd, 4, seed=seed + 1, span=6.0, sd=1.0) * 1e-2
    return P, Y0


# last iteration before the early switch (4), on it (5), between it and the momentum switch (7), on the
# momentum switch (8), on the late switch (11), after it (13), and a 60-iteration run recording at 49 and 59
@pytest.mark.parametrize("max_iter", [5, 6, 8, 9, 12, 14, 60])
def test_barnes_hut_loop_matches_reference(tmp_path, max_iter):
    P, Y0 = bh_case(160, 2, 3)
    Yr, cr = reference_run(str(tmp_path), Y0, P=P, nbody_algorithm=1, max_iter=max_iter)
    Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(max_iter=max_iter))
    assert np.array_equal(Yo, Yr)
    assert np.array_equal(co, cr)
    assert np.count_nonzero(cr) == (2 if max_iter == 60 else 1)


@pytest.mark.parametrize("nd,max_iter", [(1, 9), (2, 14), (3, 7)])
def test_no_momentum_during_exaggeration_matches_reference(tmp_path, nd, max_iter):
    P, Y0 = bh_case(140, nd, 7)
    kw = dict(no_momentum_during_exag=1, learning_rate=40.0, max_iter=max_iter)
    Yr, cr = reference_run(str(tmp_path), Y0, P=P, nbody_algorithm=1, **kw)
    Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(**kw))
    assert np.array_equal(Yo, Yr)
    assert np.array_equal(co, cr)


def test_automatic_early_exaggeration_matches_reference(tmp_path):
    # early_exag_coeff = 0: 1 / (learning_rate * max row sum of P) (tsne.cpp:391-402)
    P, Y0 = bh_case(150, 2, 11)
    kw = dict(early_exag_coeff=0.0, max_iter=4)
    Yr, cr = reference_run(str(tmp_path), Y0, P=P, nbody_algorithm=1, **kw)
    Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(**kw))
    assert np.array_equal(Yo, Yr)
    assert np.array_equal(co, cr)


def test_without_late_exaggeration_or_clip_matches_reference(tmp_path):
    P, Y0 = bh_case(150, 2, 13)
    kw = dict(start_late_exag_iter=-1, late_exag_coeff=-1.0, max_step_norm=-1.0, max_iter=9)
    Yr, cr = reference_run(str(tmp_path), Y0, P=P, nbody_algorithm=1, **kw)
    Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(**kw))
    assert np.array_equal(Yo, Yr)
    assert np.array_equal(co, cr)


def exact_case(N, seed):
    from datagen import mixture
    X, _ = mixture(N, 6, 3, seed=seed)
    Pd = orc.gaussian_perplexity_dense(X, 8.0)
    orc.oracle().orc_exact_symmetrize_normalize(Pd.reshape(-1), N)
    return Pd


@pytest.mark.parametrize("nd,df,max_iter", [(2, 1.0, 12), (1, 0.5, 6), (2, 2.0, 60), (3, 1.0, 9)])
def test_exact_loop_matches_reference(tmp_path, nd, df, max_iter):
    N = 90
    Pd = exact_case(N, 21)
    Y0 = blobs(N, nd, 3, seed=22, span=4.0, sd=1.0) * 1e-2
    kw = dict(theta=0.0, df=df, max_iter=max_iter)
    Yr, cr = reference_run(str(tmp_path), Y0, Pd=Pd, nbody_algorithm=2, **kw)
    Yo, co = orc.run_exact(Pd, Y0, **oracle_kw(**kw))
    assert np.array_equal(Yo, Yr)
    assert np.array_equal(co, cr)


def test_a_switch_moved_by_one_iteration_is_caught(tmp_path):
    """The comparison is sensitive to the schedule: moving any switch by one iteration, or dropping the KL
    correction, changes the result (a mistake planted in the oracle's settings, on the CPU)."""
    P, Y0 = bh_case(160, 2, 3)
    Yr, cr = reference_run(str(tmp_path), Y0, P=P, nbody_algorithm=1, max_iter=14)
    for shift in (dict(stop_lying_iter=4), dict(stop_lying_iter=6), dict(mom_switch_iter=7),
                  dict(mom_switch_iter=9), dict(start_late_exag_iter=10), dict(start_late_exag_iter=12)):
        Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(max_iter=14, **shift))
        assert not np.array_equal(Yo, Yr) or not np.array_equal(co, cr), shift
    # the late-exaggeration KL correction: C / late - log(late) (tsne.cpp:565-567)
    Yo, co = orc.run_bh(P, Y0, theta=BASE["theta"], **oracle_kw(max_iter=14))
    raw = orc.oracle().orc_bh_error(P[0], P[1], np.ascontiguousarray(P[2] * BASE["late_exag_coeff"]), Yo.reshape(-1),
                                    160, 2, BASE["theta"])
    assert co[13] == raw / BASE["late_exag_coeff"] - np.log(BASE["late_exag_coeff"])
