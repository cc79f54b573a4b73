"""Pins the oracle's FFT (oracle/vie_oracle.c) by the DFT definition.

The reference calls MKL DFTI (vie_fft/fft.h:69-126, vie_fft/GreenFunction.h:132-139):
forward exp(-j), backward exp(+j), both unnormalised.  MKL is not in the
image; any correct FFT agrees with the DFT to roundoff, so the oracle is
checked against brute force and numpy.fft.
"""
import numpy as np
import pytest

from oracle import oracle as O


def brute_dft(x, sign):
    n = x.size
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16])
def test_fft1d_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, +1):
        ref = brute_dft(x, sign)
        got = O.fft1d(x, sign)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [32, 64, 96, 97, 128, 192, 256])
def test_fft1d_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * np.abs(x).sum())
    assert np.allclose(O.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * np.abs(x).sum())


@pytest.mark.parametrize("shape", [(2, 3, 4), (6, 10, 12), (16, 8, 24), (64, 64, 64)])
def test_fft3d_numpy(shape):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = np.fft.fftn(a)
    got = O.fft3d(a, -1)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-14
    back = O.fft3d(got, +1) / a.size
    assert np.abs(back - a).max() < 1e-13
