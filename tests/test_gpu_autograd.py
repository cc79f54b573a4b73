"""The adjoint gradient through torch.autograd: a torch.autograd.Function whose
forward is the C-ABI forward (solve, then vie_absorption: absorbed power,
NearField's per-face term, vie_fft/NearFar.h:186-250) and whose backward is
the C-ABI backward (vie_absorption_grad: one adjoint solve with the same
operator and the fused gradient gather), checked against central finite
differences through full re-solves (64^3 design workload) and against the
oracle's FP64 gradient (16^3)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


class AbsorbedPower(torch.autograd.Function):
    """P(Re eps) for a fixed Im eps, flags, wavelength and incident field;
    the solver and the tolerance ride along as non-tensor arguments."""

    @staticmethod
    def forward(ctx, eps_re, eps_im, solver, flags, rhs, tol):
        eps = eps_re.detach().cpu().numpy() + 1j * eps_im
        solver.set_medium_eps(eps, flags)
        solver.set_rhs(rhs)
        info = solver.solve_resident(VF.BICGSTAB, tol=tol)
        assert info["converged"], info
        ctx.solver, ctx.tol = solver, tol
        return torch.tensor(solver.absorption(), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_p):
        _, gre, _, info = ctx.solver.absorption_grad(tol=ctx.tol)
        assert info["converged"], info
        return grad_p * torch.from_numpy(gre), None, None, None, None, None


def _setup(n):
    sc = G.design_case(n=n)
    eps, flags = sc.eps(), sc.flags()
    s = VF.Solver(sc.n, sc.d, sc.k0)
    s.set_medium_eps(eps, flags)
    s.green_compute()
    rhs = s.incident(*sc.rhs_angles()) * flags
    return sc, eps, flags, s, rhs


def test_autograd_gradient_matches_oracle_design16():
    sc, eps, flags, s, rhs = _setup(16)
    with s:
        x = torch.tensor(eps.real, dtype=torch.float64, requires_grad=True)
        P = AbsorbedPower.apply(x, eps.imag, s, flags, rhs, 1e-13)
        P.backward()
        J = s.solution()
    op = O.Operator(sc.n, sc.d, sc.k0, eps, flags, O.SCHED_FUSED)
    Po, gro, _, _ = O.absorption_grad(op, J, tol=1e-13)
    assert abs(P.item() - Po) <= 1e-12 * abs(Po)
    assert rel(x.grad.numpy(), gro) <= 1e-10


def test_autograd_gradient_matches_finite_differences_design64():
    """Central differences (h = 1e-4) on 20 seeded cells and along one random
    direction of all 262144 cells, each P from a fresh solve to 1e-13."""
    sc, eps, flags, s, rhs = _setup(64)
    with s:
        x = torch.tensor(eps.real, dtype=torch.float64, requires_grad=True)
        P = AbsorbedPower.apply(x, eps.imag, s, flags, rhs, 1e-13)
        P.backward()
        g = x.grad.numpy()
        h = 1e-4

        def p_at(e_re):
            with torch.no_grad():
                return AbsorbedPower.apply(torch.from_numpy(e_re), eps.imag, s, flags, rhs, 1e-13).item()

        rng = np.random.default_rng(0x5EED0005)
        gmax = np.abs(g).max()
        for _ in range(20):
            v = tuple(int(a) for a in rng.integers(0, 64, 3))
            ep, em = eps.real.copy(), eps.real.copy()
            ep[v] += h
            em[v] -= h
            fd = (p_at(ep) - p_at(em)) / (2 * h)
            assert abs(fd - g[v]) <= 1e-3 * abs(g[v]) + 1e-6 * gmax, (v, fd, g[v])
        d = rng.standard_normal(eps.shape)
        d /= np.linalg.norm(d)
        fd = (p_at(eps.real + h * d) - p_at(eps.real - h * d)) / (2 * h)
        dd = float(np.sum(g * d))
        assert abs(fd - dd) <= 1e-4 * abs(dd), (fd, dd)
