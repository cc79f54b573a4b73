"""Axis lengths without compiled pass kernels (the reference's sizing rule
N = int(Box/delta) + 2 gives any length): the grid is embedded in the next
compiled length with inactive background cells, and every result on the
caller's cells equals the oracle's at the caller's size."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.dense_ref import own_axis_flags
from tests.gpu_util import random_field, rel
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu

N_ODD = (34, 20, 27)   # 34 = 2 x 17 and 27 have no compiled kernel; 20 does
D = (0.02, 0.025, 0.03)
K0 = 2 * np.pi / 4.0


def _medium(n, seed):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.5, 6.0, n) - 1j * rng.uniform(0.0, 1.0, n)
    flags = own_axis_flags(n) * (rng.uniform(0, 1, (3,) + tuple(n)) < 0.85).astype(np.uint8)
    return eps, flags


def _solver(n, eps, flags, precision=VF.FP64):
    s = VF.Solver(n, D, K0, precision)
    s.set_medium_eps(eps, flags)
    s.green_compute()
    return s


def test_embedded_operator_matches_oracle():
    assert not VF.supported_n(34) and not VF.supported_n(27)
    eps, flags = _medium(N_ODD, 1)
    op = O.Operator(N_ODD, D, K0, eps, flags, O.SCHED_FUSED)
    x = random_field(N_ODD, 2) * flags
    with _solver(N_ODD, eps, flags) as s:
        assert rel(s.apply(x), op.apply(x)) <= 1e-12
        assert rel(s.apply(x, adjoint=True), op.apply(x, adjoint=True)) <= 1e-12
    with _solver(N_ODD, eps, flags, VF.FP32) as s:
        assert rel(s.apply(x), op.apply(x)) <= 3e-6


@pytest.mark.parametrize("precision", [VF.FP64, VF.FP32])
def test_embedded_apply_async_equals_apply(precision):
    import torch

    eps, flags = _medium(N_ODD, 3)
    x = random_field(N_ODD, 4) * flags
    dt = torch.complex128 if precision == VF.FP64 else torch.complex64
    with _solver(N_ODD, eps, flags, precision) as s:
        ref = s.apply(x)
        xt = torch.tensor(x, dtype=dt, device="cuda")
        yt = torch.empty_like(xt)
        s.apply_async(xt.data_ptr(), yt.data_ptr())
        torch.cuda.synchronize()
        got = yt.cpu().numpy()
    tol = 1e-14 if precision == VF.FP64 else 1e-6
    assert rel(got, ref) <= tol


def test_embedded_solve_and_outputs_match_oracle():
    sc = G.sphere_case(n=30, radius=9.0, eps_phys=4.0 + 0.5j)  # 30 -> compiled 32
    assert not VF.supported_n(30)
    eps, flags, mat = sc.eps(), sc.flags(), sc.material()
    op = O.Operator(sc.n, sc.d, sc.k0, eps, flags, O.SCHED_FUSED)
    b = O.incident(sc.n, sc.d, sc.k0) * flags
    ref = op.solve(b, 3, tol=1e-10)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(mat, sc.eps_table, flags)
        s.green_compute()
        assert rel(s.incident(), O.incident(sc.n, sc.d, sc.k0)) <= 1e-13
        x, info = s.solve(s.incident(), VF.BICGSTAB, tol=1e-10)
        assert abs(info["iterations"] - ref["iterations"]) <= 1
        assert rel(x, ref["x"]) <= 1e-8
        J = ref["x"]
        s.set_rhs(b)
        s.solve_resident(VF.BICGSTAB, tol=1e-10)
        xs = s.solution()
        assert abs(s.forward_scatter() - O.forward_scatter(sc.n, sc.d, sc.k0, xs)) <= 1e-11 * abs(s.forward_scatter())
        assert abs(s.absorbed(1) - O.absorbed(sc.n, sc.d, sc.k0, xs, eps, mat, 1)) <= 1e-11 * abs(s.absorbed(1))
        assert abs(s.absorption() - O.absorption(sc.n, sc.d, sc.k0, xs, eps, flags)) <= 1e-12 * s.absorption()
        rt, rp = s.rcs(0.0)
        ot, opp = O.rcs(sc.n, sc.d, sc.k0, xs, 0.0, 0.0)
        ok = np.isfinite(ot) & (ot > np.nanmax(ot) - 100)
        assert np.max(np.abs(rt[ok] - ot[ok])) < 1e-9
        assert rel(xs, J) <= 1e-8


def test_embedded_gradient_matches_oracle():
    n = (10, 9, 7)
    eps, _ = _medium(n, 5)
    flags = own_axis_flags(n)
    with _solver(n, eps, flags) as s:
        s.set_rhs(s.incident() * flags)
        s.solve_resident(VF.BICGSTAB, tol=1e-13)
        J = s.solution()
        P, gre, gim, _ = s.absorption_grad(tol=1e-13)
    op = O.Operator(n, D, K0, eps, flags, O.SCHED_FUSED)
    Po, gro, gio, _ = O.absorption_grad(op, J, tol=1e-13)
    assert gre.shape == n
    assert abs(P - Po) <= 1e-12 * abs(Po)
    assert rel(gre, gro) <= 1e-9 and rel(gim, gio) <= 1e-9


def test_embedded_grid_rejects_active_last_plane():
    eps, flags = _medium(N_ODD, 6)
    flags[0, -1, 3, 4] = 1
    with VF.Solver(N_ODD, D, K0) as s:
        with pytest.raises(VF.VieError):
            s.set_medium_eps(eps, flags)


def test_embedded_apply_of_unmasked_fields_matches_oracle():
    """Caller fields that are not zero on inactive faces or on a component's
    own-axis last plane: that plane is outside the operator on the caller's
    grid (Current_Expansion, vie_fft/CG_Iteration.h:59-63), also where the
    embedding makes it interior."""
    import torch

    n = (18, 14, 11)  # 18 and 11 run embedded (in 20 and 12)
    assert not VF.supported_n(18) and not VF.supported_n(11)
    eps, flags = _medium(n, 11)
    op = O.Operator(n, D, K0, eps, flags, O.SCHED_FUSED)
    x = random_field(n, 12)
    with _solver(n, eps, flags) as s:
        for adj in (False, True):
            assert rel(s.apply(x, adjoint=adj), op.apply(x, adjoint=adj)) <= 1e-12
        xt = torch.tensor(x, dtype=torch.complex128, device="cuda")
        yt = torch.empty_like(xt)
        s.apply_async(xt.data_ptr(), yt.data_ptr())
        torch.cuda.synchronize()
        assert rel(yt.cpu().numpy(), op.apply(x)) <= 1e-12
