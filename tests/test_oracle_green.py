"""Pins the oracle's Green's function (vie_fft/GreenFunction.h:8-141).

g is the volume integral of exp(-j k0 R)/(4 pi R) over the equivalent sphere
of the cell (diameter D with pi D^3/6 = dV).  Outside the sphere that
integral is exp(-j k0 R)(sin p - p cos p)/(k0^3 R) with p = k0 D/2; at R = 0
it is ((1 + j p) exp(-j p) - 1)/k0^2.  Both are checked here against direct
numerical quadrature, and the transform against numpy's FFT of the even
extension.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O


def sphere_integral(k0, a, R):
    """int over |r'| <= a of exp(-j k0 |R z - r'|)/(4 pi |R z - r'|) dV' (axisymmetric)."""
    def f(theta, r, part):
        dist = np.sqrt(R * R + r * r - 2 * R * r * np.cos(theta))
        v = np.exp(-1j * k0 * dist) / (4 * np.pi * dist) * 2 * np.pi * r * r * np.sin(theta)
        return v.real if part == 0 else v.imag
    re = integrate.dblquad(lambda t, r: f(t, r, 0), 0, a, 0, np.pi, epsabs=1e-14, epsrel=1e-11)[0]
    im = integrate.dblquad(lambda t, r: f(t, r, 1), 0, a, 0, np.pi, epsabs=1e-14, epsrel=1e-11)[0]
    return re + 1j * im


@pytest.mark.parametrize("d", [(0.02, 0.02, 0.02), (0.02, 0.03, 0.025)])
def test_green_offsets_match_quadrature(d):
    n = (3, 2, 2)
    k0 = 2 * np.pi / 4.0
    g = O.green_offsets(n, d, k0)
    dV = d[0] * d[1] * d[2]
    a = (3 * dV / (4 * np.pi)) ** (1 / 3)
    # self term: 1-D radial integral
    self_ref = integrate.quad(lambda r: (np.exp(-1j * k0 * r) * r).real, 0, a, epsabs=1e-16)[0] + \
        1j * integrate.quad(lambda r: (np.exp(-1j * k0 * r) * r).imag, 0, a, epsabs=1e-16)[0]
    assert abs(g[0, 0, 0] - self_ref) <= 1e-10 * abs(self_ref)
    for idx in [(1, 0, 0), (0, 1, 1), (2, 1, 2), (3, 2, 1)]:
        R = np.sqrt(sum((idx[c] * d[c]) ** 2 for c in range(3)))
        ref = sphere_integral(k0, a, R)
        assert abs(g[idx] - ref) <= 1e-8 * abs(ref), (idx, g[idx], ref)


def test_green_transform_is_fft_of_even_extension():
    n = (4, 3, 5)
    d = (0.02, 0.02, 0.02)
    k0 = 2 * np.pi / 5.5
    goff = O.green_offsets(n, d, k0)
    P = [2 * v for v in n]
    ext = np.zeros(P, complex)
    for i in range(P[0]):
        for j in range(P[1]):
            for k in range(P[2]):
                ext[i, j, k] = goff[min(i, P[0] - i), min(j, P[1] - j), min(k, P[2] - k)]
    G = O.green(n, d, k0)
    ref = np.fft.fftn(ext)
    assert np.abs(G - ref).max() <= 1e-13 * np.abs(ref).max()
    # evenness of the extension makes G~ even per axis
    assert np.abs(G - G[::-1, :, :][np.r_[-1, 0:P[0] - 1]]).max() <= 1e-12 * np.abs(G).max()
