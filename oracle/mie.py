"""Mie series for a homogeneous sphere (physics check of the cross-sections).

TEST INFRASTRUCTURE ONLY.  Restates the reference's Mie oracle
(Mie_LPR_Multilayer.m:33-185, one layer: L = 2) from textbook Mie theory
(Bohren & Huffman, ch. 4) with the extinction/scattering sums of
Mie_LPR_Multilayer.m:138-177 and the normalisation by pi r^2 (:175-177).
The reference keeps MAX = 4 terms (:30); ``nmax=None`` picks the Wiscombe
count so that more terms change nothing.

Riccati-Bessel functions are computed with our own recurrences: psi_n by
downward recurrence of the logarithmic derivative, xi_n = psi_n - j chi_n by
upward recurrence; they are cross-checked against scipy.special in the tests.
Permittivities are in the physics convention (loss Im > 0); pass the
conjugate of an engineering-sign value.
"""
from __future__ import annotations

import numpy as np


def riccati(nmax, z):
    """psi_n(z) = z j_n(z) and chi_n(z) = -z y_n(z), n = 0..nmax, complex z."""
    z = complex(z)
    # downward recurrence for D_n = psi_n'/psi_n, then psi_n upward from psi_0
    nstart = int(max(nmax, abs(z)) + 16 + 2 * abs(z) ** (1 / 3) * 4)
    D = np.zeros(nstart + 2, complex)
    for n in range(nstart, 0, -1):
        D[n - 1] = n / z - 1.0 / (D[n] + n / z)
    psi = np.zeros(nmax + 1, complex)
    psi[0] = np.sin(z)
    for n in range(1, nmax + 1):
        psi[n] = psi[n - 1] / (D[n] + n / z)
    chi = np.zeros(nmax + 1, complex)
    chi[0] = np.cos(z)
    if nmax >= 1:
        chi[1] = chi[0] / z + np.sin(z)
    for n in range(2, nmax + 1):
        chi[n] = (2 * n - 1) / z * chi[n - 1] - chi[n - 2]
    return psi, chi, D[: nmax + 1]


def coefficients(x, m, nmax):
    """Mie a_n, b_n (n = 1..nmax) for size parameter x and relative index m."""
    psi, chi, _ = riccati(nmax, x)
    xi = psi - 1j * chi                        # psi_n + j y-part: outgoing exp(+j x) (physics)
    _, _, D = riccati(nmax, m * x)
    n = np.arange(1, nmax + 1)
    psi_n, psi_m = psi[1:], psi[:-1]
    xi_n, xi_m = xi[1:], xi[:-1]
    Dn = D[1:]
    a = ((Dn / m + n / x) * psi_n - psi_m) / ((Dn / m + n / x) * xi_n - xi_m)
    b = ((Dn * m + n / x) * psi_n - psi_m) / ((Dn * m + n / x) * xi_n - xi_m)
    return a, b


def efficiencies(radius, wavelength, eps_sphere, eps_host=1.0, nmax=None):
    """(Q_ext, Q_sca, Q_abs) normalised by pi r^2 (physics-sign permittivities)."""
    nh = np.sqrt(complex(eps_host))
    k = 2 * np.pi * nh.real / wavelength
    x = k * radius
    m = np.sqrt(complex(eps_sphere)) / nh
    if m.imag < 0:
        m = -m
    if nmax is None:
        nmax = int(np.ceil(x + 4.05 * x ** (1 / 3) + 2))
        nmax = max(nmax, 4)
    a, b = coefficients(x, m, nmax)
    n = np.arange(1, nmax + 1)
    qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
    return float(qext), float(qsca), float(qext - qsca)
