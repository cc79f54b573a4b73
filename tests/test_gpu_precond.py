"""Jacobi-preconditioned BiCGSTAB on the GPU (vie_set_preconditioner).

The reference's BICGSTAB_FFT (vie_fft/CG_Iteration.h:531-776) has no
preconditioner; this option iterates on the right-preconditioned system
A (m (.) y) = b with m = 1/d_f (the diagonal face term of
vie_fft/fft.h:405-588) and returns x = m (.) y, so the stop rule still reads
the true residual ||b - A x|| / ||b||.

* Iterates: the oracle's BiCGSTAB run on the operator y -> A (m (.) y) (the
  oracle's own product, with m from the oracle's diagonal) must give the GPU's
  iterate after 8 iterations (FP64 <= 1e-10 on the roundoff-insensitive
  dielectric sphere; FP32 <= 1e-4).
* Solutions: on a layered medium like the metric's (five materials, face
  diagonals 15x apart) the preconditioned solve reaches the same solution as
  the plain one in fewer iterations, in FP64 and in the mixed solve, also
  after the permittivities change (the preconditioner is rebuilt), and from a
  warm start (the correction solve is preconditioned too).
* Everything it does not cover returns VIE_ERR_ARG.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import oracle_op, rel, scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def _sphere(eps_phys=4.0 + 0.5j):
    sc = G.sphere_case(16, radius=6.0, eps_phys=eps_phys, delta_nm=2.0)
    eps, flags = sc.eps(), sc.flags()
    b = O.incident(sc.n, sc.d, sc.k0) * flags
    return sc, eps, flags, b


def _jacobi(sc, eps, flags):
    d = O.diag(sc.n, sc.d, sc.k0, eps)
    return np.where(flags == 1, 1.0 / np.where(flags == 1, d, 1.0), 0.0)


def _oracle_precond_iterate(sc, eps, flags, b, iters):
    op = oracle_op(sc.n, sc.d, sc.k0, eps, flags)
    m = _jacobi(sc, eps, flags)
    shape = (3,) + tuple(sc.n)
    r = O.solve(3, lambda y, adj: op.apply(m * y.reshape(shape)).ravel(), b.ravel(), tol=1e-300, maxit=iters)
    return m * r["x"].reshape(shape), r["iterations"]


@pytest.mark.parametrize("precision,bound", [(VF.FP64, 1e-10), (VF.FP32, 1e-4)])
def test_preconditioned_iterates_match_oracle(precision, bound):
    sc, eps, flags, b = _sphere()
    ref, its = _oracle_precond_iterate(sc, eps, flags, b, 8)
    with scene_solver(sc, precision) as s:
        s.set_preconditioner(VF.PRECOND_JACOBI)
        x, info = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
    assert info["iterations"] == its == 8
    assert rel(x, ref) <= bound, rel(x, ref)


def _layered():
    # the metric's layer stack and sphere material in a 16 x 16 x 52 grid
    return G.layered_case(n=(16, 16, 52), band_start=2, sphere_z=30.0)


def test_preconditioned_solution_equals_plain_in_fewer_iterations():
    sc = _layered()
    with scene_solver(sc) as s:
        b = s.incident()
        x0, i0 = s.solve(b, VF.BICGSTAB, tol=1e-10, maxit=5000)
        s.set_preconditioner(VF.PRECOND_JACOBI)
        x1, i1 = s.solve(b, VF.BICGSTAB, tol=1e-10, maxit=5000)
        assert i0["converged"] and i1["converged"]
        assert i1["true_relres"] <= 2e-10 and i0["true_relres"] <= 2e-10
        assert rel(x1, x0) <= 1e-7, rel(x1, x0)
        assert i1["iterations"] < 0.8 * i0["iterations"], (i1["iterations"], i0["iterations"])
        # a permittivity step on the same geometry rebuilds d_f and with it m
        table = sc.eps_table.copy()
        table[1:] = table[1:] * (1.0 + 0.05j)
        s.set_medium(sc.material(), table, sc.flags())
        x2, i2 = s.solve(b, VF.BICGSTAB, tol=1e-10, maxit=5000)
        s.set_preconditioner(VF.PRECOND_NONE)
        x3, i3 = s.solve(b, VF.BICGSTAB, tol=1e-10, maxit=5000)
        assert i2["converged"] and i3["converged"] and i2["true_relres"] <= 2e-10
        assert rel(x2, x3) <= 1e-7, rel(x2, x3)
        assert rel(x2, x0) > 1e-3  # the medium did change


def test_off_again_gives_the_reference_iterates():
    sc, eps, flags, b = _sphere(-4.0 + 0.7j)
    with scene_solver(sc) as s:
        xa, ia = s.solve(b, VF.BICGSTAB, tol=1e-6)
        s.set_preconditioner(VF.PRECOND_JACOBI)
        s.solve(b, VF.BICGSTAB, tol=1e-6)
        s.set_preconditioner(VF.PRECOND_NONE)
        xb, ib = s.solve(b, VF.BICGSTAB, tol=1e-6)
    assert ia["iterations"] == ib["iterations"]
    assert np.array_equal(xa, xb)


def test_preconditioned_mixed_solve_matches_fp64():
    sc = _layered()
    with scene_solver(sc) as s:
        s.set_rhs_incident()
        ref = s.solve_resident(VF.BICGSTAB, tol=1e-10, maxit=5000)
        x64 = s.solution()
        q64 = s.forward_scatter()
        plain = s.solve_mixed(tol=1e-8)
        s.set_preconditioner(VF.PRECOND_JACOBI)
        pre = s.solve_mixed(tol=1e-8)
        xm = s.solution()
        qm = s.forward_scatter()
    assert ref["converged"] and plain["converged"] and pre["converged"]
    assert pre["true_relres"] < 1e-8
    assert rel(xm, x64) <= 1e-5, rel(xm, x64)
    assert abs(qm - q64) <= 1e-6 * abs(q64)
    assert pre["iterations"] < 0.8 * plain["iterations"], (pre["iterations"], plain["iterations"])


def test_preconditioned_warm_start():
    # a guess runs as the correction solve A e = b - A w, itself preconditioned
    sc = _layered()
    with scene_solver(sc) as s:
        b = s.incident()
        s.set_preconditioner(VF.PRECOND_JACOBI)
        x_ref, cold = s.solve(b, VF.BICGSTAB, tol=1e-10, maxit=5000)
        guess, _ = s.solve(b, VF.BICGSTAB, tol=1e-3, maxit=5000)
        x, warm = s.solve_x0(b, guess, VF.BICGSTAB, tol=1e-10, maxit=5000)
        assert cold["converged"] and warm["converged"] and warm["true_relres"] <= 2e-10
        assert rel(x, x_ref) <= 1e-7, rel(x, x_ref)
        assert warm["iterations"] < cold["iterations"], (warm["iterations"], cold["iterations"])
        # the mixed solve from the resident solution as its guess
        s.set_rhs(b)
        s.solve_resident(VF.BICGSTAB, tol=1e-3, maxit=5000)
        s.set_guess()
        mixed = s.solve_mixed(tol=1e-8)
        assert mixed["converged"] and mixed["true_relres"] < 1e-8
        assert rel(s.solution(), x_ref) <= 1e-6, rel(s.solution(), x_ref)


def _status(fn):
    try:
        fn()
    except VF.VieError as e:
        return e.status
    return VF.OK


def test_what_it_does_not_cover_is_rejected():
    sc, eps, flags, b = _sphere()
    with scene_solver(sc) as s:
        assert _status(lambda: s.set_preconditioner(7)) == VF.ERR_ARG
        s.set_preconditioner(VF.PRECOND_JACOBI)
        for method in (VF.CG, VF.BICG, VF.CORS, VF.BICORSTAB):
            assert _status(lambda: s.solve(b, method, tol=1e-6)) == VF.ERR_ARG
        s.batch_reserve(2)
        s.batch_set_rhs(0, b)
        s.batch_set_rhs(1, 2 * b)
        assert _status(lambda: s.batch_solve(2, VF.BICGSTAB, tol=1e-6)) == VF.ERR_ARG
        assert _status(lambda: s.batch_solve(2, VF.BICGSTAB, tol=1e-6, mixed=True)) == VF.ERR_ARG
        # the context stays usable
        s.set_preconditioner(VF.PRECOND_NONE)
        x, info = s.solve(b, VF.BICGSTAB, tol=1e-6)
        assert info["converged"]


def test_vanishing_diagonal_is_rejected():
    # eps_a + eps_b = 2 makes 1/(eps_a - 1) + 1/(eps_b - 1) = 0 on the faces
    # between the two materials
    n = (8, 8, 8)
    sc = G.sphere_case(8, radius=3.0, eps_phys=3.0, delta_nm=2.0)
    eps = np.full(n, G.phys_to_eng(3.0), np.complex128)
    eps[4:] = G.phys_to_eng(-1.0)
    flags = np.ones((3,) + n, np.uint8)
    flags[0, -1] = 0
    flags[1, :, -1] = 0
    flags[2, :, :, -1] = 0
    s = VF.Solver(n, sc.d, sc.k0, VF.FP64)
    try:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        b = O.incident(n, sc.d, sc.k0) * flags
        s.set_preconditioner(VF.PRECOND_JACOBI)
        assert _status(lambda: s.solve(b, VF.BICGSTAB, tol=1e-6)) == VF.ERR_ARG
    finally:
        s.close()
