"""Pins the oracle's FFT (the stand-in for FFTW on the reference path) against the DFT definition,
numpy's FFT and the algebraic identities the convolution step relies on."""
import numpy as np
import pytest

import orc


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 12, 30, 49, 50, 64, 105, 150, 154, 300, 2 * 7 * 11 * 3])
def test_c2c_matches_dft_definition(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    for inv in (False, True):
        got = orc.fft_c2c(x, inverse=inv)
        want = orc.dft_naive(x, inverse=inv)
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.abs(want).max()) * np.log2(n + 1)


@pytest.mark.parametrize("n", [6, 150, 1050, 3600, 601 * 6])
def test_c2c_matches_numpy_and_round_trips(n):
    rng = np.random.default_rng(7 + n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    f = orc.fft_c2c(x)
    assert np.allclose(f, np.fft.fft(x), rtol=0, atol=1e-11 * np.abs(f).max())
    back = orc.fft_c2c(f, inverse=True) / n
    assert np.allclose(back, x, rtol=0, atol=1e-13 * np.abs(x).max() * np.log2(n))


@pytest.mark.parametrize("shape", [(6, 6), (150, 150), (12, 30), (300, 300)])
def test_r2c_c2r_2d(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape)
    z = orc.fft_r2c_2d(a)
    assert z.shape == (shape[0], shape[1] // 2 + 1)
    assert np.allclose(z, np.fft.rfft2(a), rtol=0, atol=1e-11 * np.abs(z).max())
    back = orc.fft_c2r_2d(z, shape[1]) / (shape[0] * shape[1])
    assert np.allclose(back, a, rtol=0, atol=1e-13 * np.log2(a.size))


def test_c2r_ignores_imaginary_part_of_self_conjugate_bins():
    # FFTW c2r semantics: a Hermitian half-spectrum's self-conjugate bins are taken as real
    rng = np.random.default_rng(3)
    a = rng.normal(size=(8, 10))
    z = orc.fft_r2c_2d(a)
    z2 = z.copy()
    z2[0, 0] += 0.5j
    z2[0, 5] += 0.25j
    assert np.allclose(orc.fft_c2r_2d(z2, 10), orc.fft_c2r_2d(z, 10), atol=1e-12)
