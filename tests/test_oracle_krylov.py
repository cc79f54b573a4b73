"""Pins the oracle's five Krylov loops (vie_fft/CG_Iteration.h:176-1388).

* BiCGSTAB is the netlib template the reference cites (:7) with shadow
  residual b; scipy.sparse.linalg.bicgstab implements the same recurrences,
  so the iterates must agree after every fixed number of iterations.
* CG_FFT is CG on the normal equations; scipy's CG on A^H A x = A^H b has the
  same iterates.
* BICG_FFT with unconjugated products is COCG, which equals CG for a real
  symmetric positive definite matrix.
* CORS and BiCOR-STAB (Jing et al., PIER 99, 2009) have no library
  counterpart: they are pinned by convergence to the dense solution and by
  finite termination on a matrix with three distinct eigenvalues (a dropped
  term or a wrong sign in a recurrence destroys both).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import oracle as O


def _cplx_matrix(n, seed, sym=False):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    if sym:
        A = 0.5 * (A + A.T)
    return A + (3.0 + 1.0j) * np.eye(n)


def _mv(A):
    return lambda x, adj: (A.conj().T @ x) if adj else (A @ x)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_bicgstab_iterates_match_scipy(k):
    A = _cplx_matrix(40, 1)
    b = np.random.default_rng(2).standard_normal(40) + 0j
    ours = O.solve(3, _mv(A), b, tol=1e-300, maxit=k)
    ref, _ = sla.bicgstab(A, b, rtol=0.0, atol=1e-300, maxiter=k)
    assert ours["iterations"] == k
    assert np.linalg.norm(ours["x"] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("k", [1, 2, 4, 7])
def test_cgnr_iterates_match_scipy_cg_on_normal_equations(k):
    A = _cplx_matrix(30, 3)
    b = np.random.default_rng(4).standard_normal(30) + 1j * np.random.default_rng(5).standard_normal(30)
    ours = O.solve(1, _mv(A), b, tol=1e-300, maxit=k)
    N = A.conj().T @ A
    ref, _ = sla.cg(N, A.conj().T @ b, rtol=0.0, atol=1e-300, maxiter=k)
    assert np.linalg.norm(ours["x"] - ref) <= 1e-11 * np.linalg.norm(ref)


@pytest.mark.parametrize("k", [1, 3, 6])
def test_bicg_is_cg_on_real_spd(k):
    rng = np.random.default_rng(6)
    B = rng.standard_normal((30, 30))
    A = B @ B.T / 30 + np.eye(30)
    b = rng.standard_normal(30) + 0j
    ours = O.solve(2, _mv(A.astype(complex)), b, tol=1e-300, maxit=k)
    ref, _ = sla.cg(A, b.real, rtol=0.0, atol=1e-300, maxiter=k)
    assert np.linalg.norm(ours["x"] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_converges_to_dense_solution(method):
    A = _cplx_matrix(50, 7, sym=True)   # complex symmetric like the VIE operator
    b = np.random.default_rng(8).standard_normal(50) + 1j * np.random.default_rng(9).standard_normal(50)
    tol = 1e-20 if method in (1, 2) else 1e-11  # CG and BiCG stop on squared quantities
    out = O.solve(method, _mv(A), b, tol=tol, maxit=500)
    assert out["converged"]
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(out["x"] - ref) <= 1e-8 * np.linalg.norm(ref)


@pytest.mark.parametrize("method", [2, 3, 4, 5])
def test_finite_termination_three_eigenvalues(method):
    """A = diag with 3 distinct values: exact arithmetic terminates in 3 steps."""
    lam = np.array([1.0 + 0.5j, 2.0 - 0.3j, 3.5 + 0.1j])
    d = np.repeat(lam, 7)
    b = np.random.default_rng(10).standard_normal(d.size) + 0j
    out = O.solve(method, lambda x, adj: (np.conj(d) * x) if adj else d * x, b, tol=1e-10, maxit=20)
    assert out["converged"] and out["iterations"] <= 3


def test_stop_rules_follow_the_reference():
    A = _cplx_matrix(40, 11, sym=True)
    b = np.random.default_rng(12).standard_normal(40) + 0j
    for method in (3, 4, 5):
        out = O.solve(method, _mv(A), b, tol=1e-6, maxit=200)
        rel = np.linalg.norm(b - A @ out["x"]) / np.linalg.norm(b)
        assert out["err"] < 1e-6 and abs(out["err"] - rel) <= 1e-9
    cg = O.solve(1, _mv(A), b, tol=1e-6, maxit=200)      # squared ratio
    rel = np.linalg.norm(b - A @ cg["x"]) / np.linalg.norm(b)
    assert cg["err"] < 1e-6 and abs(cg["err"] - rel ** 2) <= 1e-9
    bi = O.solve(2, _mv(A), b, tol=1e-6, maxit=200)      # |r^T r| / |b^T b|
    r = b - A @ bi["x"]
    assert bi["err"] < 1e-6 and abs(bi["err"] - abs(r @ r) / abs(b @ b)) <= 1e-9
