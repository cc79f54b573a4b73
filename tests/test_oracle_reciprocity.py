"""Reciprocity of the oracle's VIE solutions.  For plane waves from directions
k1, k2 with polarisations e1, e2 (Incident_Wave, vie_fft/GreenFunction.h:144-248)
and the far-field sums N (FarField1/2, vie_fft/NearFar.h:59-96), the solution
J1 of the first wave radiates towards -k2 what J2 radiates towards -k1:

    e2 . N[J1](-k2) = e1 . N[J2](-k1),

because both sides are b2^T A^{-1} b1 and the operator is complex symmetric
on cubic cells (A^T = A).  The discrete sides differ only by the incident
field's pulse integral against the far-field sums' point currents, (k0 delta)^2
small here.  A wrong incidence angle, polarisation sign or far-field phase
convention breaks it at O(1)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_energy import far_sums
from vie_fft_amd import geometry as G


def polarisation(theta, phi, psi):
    """The incident field's unit polarisation (vie_fft/GreenFunction.h:196-222)."""
    t, p, q = np.radians([theta, phi, psi])
    return np.array([np.cos(t) * np.cos(p) * np.cos(q) - np.sin(p) * np.sin(q),
                     np.cos(t) * np.sin(p) * np.cos(q) + np.cos(p) * np.sin(q),
                     -np.sin(t) * np.cos(q)])


def backward(theta, phi):
    """(theta, phi) in radians of the direction opposite to the incidence (theta, phi) in degrees."""
    t, p = np.radians([theta, phi])
    k = -np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
    return np.arccos(np.clip(k[2], -1.0, 1.0)), np.arctan2(k[1], k[0])


A1 = (20.0, 30.0, 10.0)
A2 = (50.0, 200.0, 70.0)


def reciprocity_pair(n, d, k0, J1, J2):
    t2, p2 = backward(*A2[:2])
    t1, p1 = backward(*A1[:2])
    u = polarisation(*A2) @ far_sums(n, d, k0, J1, np.array([t2]), np.array([p2]))[0]
    v = polarisation(*A1) @ far_sums(n, d, k0, J2, np.array([t1]), np.array([p1]))[0]
    return u, v


def test_polarisation_is_transverse_and_unit():
    for a in (A1, A2, (0.0, 0.0, 0.0), (90.0, 45.0, 30.0)):
        e = polarisation(*a)
        t, p = np.radians(a[:2])
        k = np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
        assert abs(np.linalg.norm(e) - 1) < 1e-14 and abs(e @ k) < 1e-14


@pytest.mark.parametrize("case", ["sphere", "layered"])
def test_far_fields_are_reciprocal(case):
    sc = (G.sphere_case(12, radius=4.0, eps_phys=3.0 + 0.4j, delta_nm=4.0) if case == "sphere"
          else G.layered_case(n=(12, 12, 20), band_start=1, sphere_z=12.0))
    n, d, k0 = sc.n, sc.d, sc.k0
    eps, flags = sc.eps(), sc.flags()
    op = O.Operator(n, d, k0, eps, flags, O.SCHED_FUSED)
    J = []
    for a in (A1, A2):
        b = O.incident(n, d, k0, *a) * flags
        r = O.solve(3, lambda x, adj: op.apply(x.reshape((3,) + tuple(n)), adjoint=adj).ravel(), b.ravel(),
                    tol=1e-10, maxit=5000)
        assert r["converged"]
        J.append(r["x"].reshape((3,) + tuple(n)))
    u, v = reciprocity_pair(n, d, k0, J[0], J[1])
    assert abs(u - v) <= 1e-4 * abs(u), (u, v)
    # and the two are not trivially equal: swapping the polarisations breaks it
    t1, p1 = backward(*A1[:2])
    w = polarisation(A1[0], A1[1], A1[2] + 90.0) @ far_sums(n, d, k0, J[1], np.array([t1]), np.array([p1]))[0]
    assert abs(u - w) > 1e-2 * abs(u)
