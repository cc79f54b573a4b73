"""Pins the oracle's absorbed-power objective and its adjoint eps gradient.

The objective's per-face weight is checked against the oracle's restatement
of NearField's Ene_Metal sum (vie_fft/NearFar.h:186-250) on a field that
lives inside that sum's i,j,k < N-1 face set; the gradient is checked
against central finite differences of the objective through dense solves of
the real-space operator assembly (no FFT, no Krylov), on cubic and
non-cubic cells.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.dense_ref import dense_operator, own_axis_flags

K0 = 2 * np.pi / 4.0


def _case(n, seed):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.5, 6.0, n) - 1j * rng.uniform(0.05, 0.8, n)
    flags = own_axis_flags(n) * (rng.uniform(0, 1, (3,) + tuple(n)) < 0.85).astype(np.uint8)
    return eps, flags


def _dense_solve(n, d, eps, flags, b):
    A = dense_operator(n, d, K0, eps, flags)
    keep = flags.reshape(-1).astype(bool)
    x = np.zeros(A.shape[0], complex)
    x[keep] = np.linalg.solve(A[np.ix_(keep, keep)], b.reshape(-1)[keep])
    return x.reshape((3,) + tuple(n))


def _rhs(n, flags, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((3,) + tuple(n)) + 1j * rng.standard_normal((3,) + tuple(n))) * flags


def test_weight_matches_nearfield_formula():
    """All cells one material and J zero on faces with any index N-1: the
    active-face objective equals the Ene_Metal-style sum exactly."""
    n, d = (5, 4, 6), (0.02, 0.03, 0.025)
    rng = np.random.default_rng(7)
    eps = rng.uniform(1.5, 6.0, n) - 1j * rng.uniform(0.05, 0.8, n)
    flags = own_axis_flags(n)
    J = rng.standard_normal((3,) + n) + 1j * rng.standard_normal((3,) + n)
    J[:, -1, :, :] = 0
    J[:, :, -1, :] = 0
    J[:, :, :, -1] = 0
    mat = np.ones(n, np.uint8)
    p_new = O.absorption(n, d, K0, J, eps, flags)
    p_ref = O.absorbed(n, d, K0, J, eps, mat, 1)
    assert p_ref > 0
    assert abs(p_new - p_ref) <= 1e-13 * abs(p_ref)


def test_lossless_medium_absorbs_nothing():
    n, d = (4, 4, 4), (0.02,) * 3
    eps = np.full(n, 3.0 + 0j)
    flags = own_axis_flags(n)
    J = _rhs(n, flags, 1)
    assert abs(O.absorption(n, d, K0, J, eps, flags)) < 1e-12


@pytest.mark.parametrize("n,d", [((4, 5, 3), (0.02, 0.02, 0.02)), ((3, 4, 5), (0.02, 0.03, 0.025))])
def test_gradient_matches_finite_differences(n, d):
    eps, flags = _case(n, 11)
    b = _rhs(n, flags, 12)

    def P(e):
        return O.absorption(n, d, K0, _dense_solve(n, d, e, flags, b), e, flags)

    J = _dense_solve(n, d, eps, flags, b)
    op = O.Operator(n, d, K0, eps, flags, O.SCHED_FUSED)
    p0, gre, gim, info = O.absorption_grad(op, J, method=3, tol=1e-14)
    assert abs(p0 - P(eps)) <= 1e-12 * abs(p0)
    rng = np.random.default_rng(13)
    h = 1e-6
    for _ in range(8):
        v = tuple(rng.integers(0, m) for m in n)
        for part, g in ((1.0, gre), (1j, gim)):
            ep, em = eps.copy(), eps.copy()
            ep[v] += h * part
            em[v] -= h * part
            fd = (P(ep) - P(em)) / (2 * h)
            assert abs(fd - g[v]) <= 1e-6 * np.max(np.abs(g)), (v, part, fd, g[v])
