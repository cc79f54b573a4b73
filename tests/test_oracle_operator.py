"""Pins the oracle's operator (both schedules) against the dense real-space
assembly, the adjoint against the dense conjugate transpose, and the
symmetry that makes BiCG/CORS applicable (vie_fft/CG_Iteration.h:402-406)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.dense_ref import dense_operator, diag_terms, own_axis_flags

CASES = [((3, 4, 5), (0.02, 0.02, 0.02)), ((4, 3, 3), (0.02, 0.03, 0.025)), ((2, 2, 3), (0.05, 0.05, 0.05))]


def _medium(n, seed):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.5, 5, n) - 1j * rng.uniform(0, 1, n)
    flags = (rng.uniform(0, 1, (3,) + tuple(n)) < 0.8).astype(np.uint8)
    return eps, flags


@pytest.mark.parametrize("n,d", CASES)
@pytest.mark.parametrize("sched", [O.SCHED_REFERENCE, O.SCHED_FUSED])
def test_operator_matches_dense(n, d, sched):
    k0 = 2 * np.pi / 4.0
    eps, flags = _medium(n, 1)
    A = dense_operator(n, d, k0, eps, flags)
    op = O.Operator(n, d, k0, eps, flags, sched)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3,) + n) + 1j * rng.standard_normal((3,) + n)
    ref = (A @ x.reshape(-1)).reshape(x.shape)
    got = op.apply(x)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("n,d", CASES)
@pytest.mark.parametrize("sched", [O.SCHED_REFERENCE, O.SCHED_FUSED])
def test_adjoint_matches_dense(n, d, sched):
    """With every usable face active the _Conjugate path is conj(A)^T."""
    k0 = 2 * np.pi / 4.0
    eps, _ = _medium(n, 3)
    flags = own_axis_flags(n)
    A = dense_operator(n, d, k0, eps, flags)
    op = O.Operator(n, d, k0, eps, flags, sched)
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((3,) + n) + 1j * rng.standard_normal((3,) + n)) * flags
    ref = (A.conj().T @ x.reshape(-1)).reshape(x.shape) * flags
    got = op.apply(x, adjoint=True)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("n,d", CASES)
def test_scaled_operator_is_complex_symmetric(n, d):
    """diag(1/delta_f) A is symmetric on the active unknowns (A^T = A for cubes)."""
    k0 = 2 * np.pi / 4.0
    eps, _ = _medium(n, 5)
    flags = own_axis_flags(n)
    A = dense_operator(n, d, k0, eps, flags)
    keep = flags.reshape(-1).astype(bool)
    V = int(np.prod(n))
    Ak = (np.repeat(1 / np.array(d), V)[:, None] * A)[np.ix_(keep, keep)]
    assert np.linalg.norm(Ak - Ak.T) <= 1e-13 * np.linalg.norm(Ak)


def test_schedules_agree_and_are_linear():
    n, d, k0 = (6, 5, 4), (0.02, 0.02, 0.02), 2 * np.pi / 5.0
    eps, flags = _medium(n, 6)
    a = O.Operator(n, d, k0, eps, flags, O.SCHED_REFERENCE)
    b = O.Operator(n, d, k0, eps, flags, O.SCHED_FUSED)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((3,) + n) + 1j * rng.standard_normal((3,) + n)
    y = rng.standard_normal((3,) + n) + 1j * rng.standard_normal((3,) + n)
    for adj in (False, True):
        ya, yb = a.apply(x, adj), b.apply(x, adj)
        assert np.linalg.norm(ya - yb) <= 1e-13 * np.linalg.norm(ya)
    lin = b.apply(2.5 * x - 1j * y) - (2.5 * b.apply(x) - 1j * b.apply(y))
    assert np.linalg.norm(lin) <= 1e-13 * np.linalg.norm(b.apply(x))


def test_diag_matches_formula():
    n, d, k0 = (3, 4, 2), (0.02, 0.03, 0.04), 2.0
    eps, _ = _medium(n, 8)
    assert np.allclose(O.diag(n, d, k0, eps), diag_terms(n, d, k0, eps), rtol=1e-15, atol=0)
    assert np.allclose(O.diag(n, d, k0, eps, adjoint=True), np.conj(diag_terms(n, d, k0, eps)), rtol=1e-15, atol=0)
