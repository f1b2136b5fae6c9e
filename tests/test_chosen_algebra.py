"""The product's cancellation-free formulation (node-relative charges, odd kernels, Z by Parseval,
2^a 3^b padding) computes the same interpolated quantities as the reference's charge expansion:
checked in FP64 against the oracle for all four gradient variants, before any GPU code relies on it."""
import numpy as np
import pytest

import orc
from chosen_model import chosen_gradient, fft_size
from datagen import blobs, random_csr


def test_fft_size():
    assert fft_size(2 * 525 - 1) == 1152
    assert fft_size(2 * 1800 - 1) == 3888
    assert fft_size(2 * 150 - 1) == 324
    assert fft_size(2 * 240 - 1) == 512  # short lines: 512 = 16 x 16 x 2 (3 stages) over 486 (4)
    assert fft_size(2 * 105 - 1) == 256  # 16 x 16 over 216 = 8 x 27 (3 stages)
    assert fft_size(2 * 450 - 1) == 972  # longer lines: the smallest
    for n in range(1, 5000, 3):
        pw = 1
        while pw < n:
            pw *= 2
        assert n <= fft_size(n) <= pw
    for n in range(1, 3000, 7):
        m = fft_size(n)
        assert m >= n
        while m % 2 == 0:
            m //= 2
        while m % 3 == 0:
            m //= 3
        assert m == 1


@pytest.mark.parametrize("D,df", [(2, 1.0), (2, 0.5), (2, 3.0), (1, 1.0), (1, 0.5)])
def test_chosen_formulation_equals_reference_in_fp64(D, df):
    N = 2500
    P = random_csr(N, 10, seed=3)
    Y = blobs(N, D, 8, seed=4, span=70.0, sd=2.0)
    g, Z = orc.gradient(P, Y, df=df)
    gc, Zc = chosen_gradient(P, Y, df=df)
    assert abs(Z - Zc) / abs(Z) < 1e-12
    assert np.linalg.norm(g - gc) / np.linalg.norm(g) < 1e-12
