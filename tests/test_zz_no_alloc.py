"""The KL entry point that allocates nothing (fitsne_evaluate_error_fft_noalloc) against the allocating one.

Both run the same kernels with the same launch shape on the same values, so C must be bitwise equal. Two
calls in a row, the second with a smaller input, check that nothing the first call left in the shared
device array leaks into the second; a size beyond the array is refused on the host.  Collected last, so
that a failure here cannot hide the other tests.
"""
import numpy as np
import pytest

import fitsne_gpu as fg
import orc
from datagen import blobs, random_csr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("D,df", [(2, 1.0), (1, 0.6)])
def test_noalloc_kl_equals_the_allocating_entry_point(D, df):
    cases = [(9000, 12, 3), (2500, 7, 5)]  # the second call is smaller than the first
    for N, per_row, seed in cases:
        say(f"[gpu] evaluate_error_fft_noalloc N={N} D={D} df={df}")
        P = random_csr(N, per_row, seed=seed)
        Y = blobs(N, D, 6, seed=seed + 1, span=30.0)
        if D == 1:
            Y = Y[:, 0].copy()
        zq = 1234.5 + N
        got = fg.evaluate_error_fft_noalloc(P, Y, zq, df=df)
        ref = fg.evaluate_error_fft(P, Y, zq, df=df)
        assert got == ref, (N, got, ref)
        want = orc.kl(P, Y if Y.ndim == 2 else Y[:, None], zq, df=df)
        assert abs(got - want) <= 1e-12 * abs(want)


def test_noalloc_kl_refuses_inputs_beyond_its_scratch():
    N = 3_000_000  # 8 N D bytes of Y alone exceed the 32 MiB array
    row = np.zeros(N + 1, dtype=np.uint32)
    col = np.zeros(1, dtype=np.uint32)
    val = np.zeros(1)
    Y = np.zeros((N, 2))
    with pytest.raises(fg.FitsneError) as e:
        fg.evaluate_error_fft_noalloc((row, col, val), Y, 1.0)
    assert e.value.code == fg.E_ARG
    assert "scratch" in str(e.value)
