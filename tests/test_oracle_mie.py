"""Mie series (restating Mie_LPR_Multilayer.m for one homogeneous layer) and
the physics check of the VIE cross-sections against it.

The Riccati-Bessel recurrences are pinned against scipy.special; the series
against the Rayleigh limit and energy conservation; the CPU VIE path
(FarField1 / NearField restated from vie_fft/NearFar.h) against Mie within
the staircase error, which must shrink under refinement.
"""
import numpy as np
import pytest
import scipy.special as sp

from oracle import mie
from oracle import oracle as O
from vie_fft_amd import geometry as G

Z0 = 120 * np.pi


@pytest.mark.parametrize("z", [0.3 + 0.0j, 2.5 + 0.1j, 1.2 + 1.9j, 7.0 - 0.4j])
def test_riccati_matches_scipy(z):
    nmax = 12
    psi, chi, _ = mie.riccati(nmax, z)
    n = np.arange(nmax + 1)
    assert np.allclose(psi, z * sp.spherical_jn(n, z), rtol=1e-11, atol=1e-14)
    assert np.allclose(chi, -z * sp.spherical_yn(n, z), rtol=1e-11, atol=1e-14)


def test_rayleigh_limit():
    e = -4.0 + 0.7j
    x = 2 * np.pi * 0.5 / 400
    qabs = mie.efficiencies(0.5, 400, e)[2]
    assert abs(qabs - 4 * x * ((e - 1) / (e + 2)).imag) <= 1e-3 * qabs


def test_lossless_sphere_conserves_energy():
    qe, qs, qa = mie.efficiencies(60, 500, 2.25 + 0j)
    assert abs(qa) <= 1e-12 * qe and qe > 0


def test_reference_term_count_is_enough():
    """Mie_LPR_Multilayer.m:30 keeps MAX = 4 terms; at x ~ 0.3 more change nothing."""
    full = mie.efficiencies(20, 400, -4 + 0.7j)
    four = mie.efficiencies(20, 400, -4 + 0.7j, nmax=4)
    assert np.allclose(full, four, rtol=1e-9)


def _vie_q(n, R, eps_phys):
    sc = G.sphere_case(n, radius=R, eps_phys=eps_phys, delta_nm=20.0 / R)
    mat, fl, eps = sc.material(), sc.flags(), sc.eps()
    op = O.Operator(sc.n, sc.d, sc.k0, eps, fl, O.SCHED_FUSED)
    b = O.incident(sc.n, sc.d, sc.k0) * fl
    out = op.solve(b, 3, tol=1e-8)
    J = out["x"]
    qe = O.forward_scatter(sc.n, sc.d, sc.k0, J) / sc.area
    qa = O.absorbed(sc.n, sc.d, sc.k0, J, eps, mat, 1) * Z0 / sc.area
    return qe, qa


def test_vie_converges_to_mie_for_a_dielectric_sphere():
    eps = 2.25 + 0.2j
    qe_m, _, qa_m = mie.efficiencies(20, 400, eps)
    errs = []
    for n, R in [(24, 7.5), (32, 10.0)]:
        qe, qa = _vie_q(n, R, eps)
        errs.append((abs(qe - qe_m) / qe_m, abs(qa - qa_m) / qa_m))
    assert errs[1][0] < errs[0][0] and errs[1][1] < errs[0][1]     # refinement reduces the error
    assert errs[1][0] < 0.1 and errs[1][1] < 0.1                   # staircase error at R = 10 cells
