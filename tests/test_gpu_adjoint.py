"""GPU absorbed power and its adjoint permittivity gradient (the design-gradient
workload: 64^3 grid, 5 nm cells, 600 nm, every cell a design variable with
eps = U[1, 12] - 0.1j) against the oracle and against central finite
differences of the objective through full re-solves; the mixed-precision
adjoint (vie_absorption_grad_mixed) against the FP64 one."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def _solver(sc, eps, flags):
    s = VF.Solver(sc.n, sc.d, sc.k0)
    s.set_medium_eps(eps, flags)
    s.green_compute()
    b = s.incident(*sc.rhs_angles()) * flags
    s.set_rhs(b)
    return s, b


def _noncubic_case():
    n, d = (10, 12, 8), (0.05, 0.04, 0.06)
    rng = np.random.default_rng(21)
    eps = rng.uniform(1.5, 8.0, n) - 1j * rng.uniform(0.05, 0.5, n)
    flags = np.ones((3,) + n, np.uint8)
    flags[0, -1], flags[1, :, -1], flags[2, :, :, -1] = 0, 0, 0
    return n, d, 2 * np.pi / 6.0, eps, flags


def test_absorption_and_gradient_match_oracle_design16():
    sc = G.design_case(n=16)
    eps, flags = sc.eps(), sc.flags()
    with _solver(sc, eps, flags)[0] as s:
        info = s.solve_resident(VF.BICGSTAB, tol=1e-13)
        J = s.solution()
        P_alone = s.absorption()
        P, gre, gim, ai = s.absorption_grad(tol=1e-13)
    assert info["converged"] and ai["converged"]
    assert P == P_alone
    assert abs(P - O.absorption(sc.n, sc.d, sc.k0, J, eps, flags)) <= 1e-13 * P
    op = O.Operator(sc.n, sc.d, sc.k0, eps, flags, O.SCHED_FUSED)
    Po, gro, gio, _ = O.absorption_grad(op, J, tol=1e-13)
    assert rel(gre, gro) <= 1e-10 and rel(gim, gio) <= 1e-10


def test_gradient_matches_oracle_noncubic_cells():
    """A^T = diag(delta) A diag(1/delta) scaling of the adjoint right-hand side."""
    n, d, k0, eps, flags = _noncubic_case()
    with VF.Solver(n, d, k0) as s:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        s.set_rhs(s.incident() * flags)
        s.solve_resident(VF.BICGSTAB, tol=1e-13)
        J = s.solution()
        P, gre, gim, _ = s.absorption_grad(tol=1e-13)
    op = O.Operator(n, d, k0, eps, flags, O.SCHED_FUSED)
    Po, gro, gio, _ = O.absorption_grad(op, J, tol=1e-13)
    assert abs(P - Po) <= 1e-13 * abs(Po)
    assert rel(gre, gro) <= 1e-10 and rel(gim, gio) <= 1e-10


def test_gradient_matches_finite_differences_design64():
    """dP/dRe(eps_v) on 20 seeded cells against (P(eps + h) - P(eps - h)) / 2h,
    h = 1e-4, every P from a fresh BiCGSTAB solve to 1e-13."""
    sc = G.design_case(n=64)
    eps, flags = sc.eps(), sc.flags()
    s, b = _solver(sc, eps, flags)
    with s:
        info = s.solve_resident(VF.BICGSTAB, tol=1e-13)
        P, gre, gim, ai = s.absorption_grad(tol=1e-13)
        assert info["converged"] and ai["converged"] and P > 0
        rng = np.random.default_rng(0x5EED0005)
        h = 1e-4
        gmax = np.abs(gre).max()
        for _ in range(20):
            v = tuple(int(a) for a in rng.integers(0, 64, 3))
            vals = []
            for sgn in (1.0, -1.0):
                e = eps.copy()
                e[v] += sgn * h
                s.set_medium_eps(e, flags)
                s.set_rhs(b)
                assert s.solve_resident(VF.BICGSTAB, tol=1e-13)["converged"]
                vals.append(s.absorption())
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - gre[v]) <= 1e-3 * abs(gre[v]) + 1e-6 * gmax, (v, fd, gre[v])


@pytest.mark.parametrize("case", ["design16", "noncubic"])
def test_mixed_adjoint_equals_fp64_adjoint(case):
    """vie_absorption_grad_mixed: the adjoint system by the mixed-precision
    BiCGSTAB (FP32 iterations, FP64 residual replacements) to the same true
    residual gives the FP64 gradient."""
    if case == "design16":
        sc = G.design_case(n=16)
        n, d, k0, eps, flags = sc.n, sc.d, sc.k0, sc.eps(), sc.flags()
    else:
        n, d, k0, eps, flags = _noncubic_case()
    with VF.Solver(n, d, k0) as s:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        s.set_rhs(s.incident() * flags)
        s.solve_resident(VF.BICGSTAB, tol=1e-13)
        J = s.solution()
        P, gre, gim, a64 = s.absorption_grad(tol=1e-12)
        Pm, grm, gim_m, am = s.absorption_grad(tol=1e-11, mixed=True)
        assert np.array_equal(s.solution(), J)  # the forward solution stays resident
    assert a64["converged"] and am["converged"] and am["true_relres"] < 1e-11 and am["outer"] >= 1
    assert Pm == P
    assert rel(grm, gre) <= 1e-8 and rel(gim_m, gim) <= 1e-8


def test_mixed_forward_and_adjoint_design64():
    """The design workload with both solves mixed (to 1e-10): the FP64 gradient within 1e-7."""
    sc = G.design_case(n=64)
    eps, flags = sc.eps(), sc.flags()
    s, b = _solver(sc, eps, flags)
    with s:
        s.solve_resident(VF.BICGSTAB, tol=1e-10)
        P, gre, gim, _ = s.absorption_grad(tol=1e-10)
        fwd = s.solve_mixed(tol=1e-10)
        Pm, grm, gim_m, am = s.absorption_grad(tol=1e-10, mixed=True)
    assert fwd["converged"] and am["converged"] and am["true_relres"] < 1e-10
    assert abs(Pm - P) <= 1e-9 * P
    assert rel(grm, gre) <= 1e-7 and rel(gim_m, gim) <= 1e-7


def test_gradient_needs_a_solution():
    sc = G.design_case(n=8)
    with _solver(sc, sc.eps(), sc.flags())[0] as s:
        with pytest.raises(VF.VieError):
            s.absorption_grad()
        with pytest.raises(VF.VieError):
            s.absorption_grad(mixed=True)
        with pytest.raises(VF.VieError):
            s.absorption()


def test_mixed_gradient_needs_an_fp64_context():
    sc = G.design_case(n=8)
    eps, flags = sc.eps(), sc.flags()
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP32) as s:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        s.set_rhs(s.incident(*sc.rhs_angles()) * flags)
        s.solve_resident(VF.BICGSTAB, tol=1e-5)
        with pytest.raises(VF.VieError) as e:
            s.absorption_grad(mixed=True)
        assert e.value.status == VF.ERR_ARG
        # the plain FP32 gradient still works on it
        P, gre, gim, info = s.absorption_grad(tol=1e-5)
        assert P > 0 and info["converged"]
