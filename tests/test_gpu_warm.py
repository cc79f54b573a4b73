"""Warm starts (vie_set_guess, vie_solve_x0): the reference always starts
from x = 0 (vie_fft/Model.h:189-191); a guess must leave the stop rule
||b - A x|| / ||b|| < tol of vie_fft/CG_Iteration.h:727-740 unchanged."""
import numpy as np
import pytest

from tests.gpu_util import rel, scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def _scene(lam=450.0):
    return G.sphere_case(24, radius=7.0, eps_phys=complex(G.drude_eps_phys(lam)), delta_nm=2.0, wavelength_nm=lam)


@pytest.mark.parametrize("method", [VF.BICGSTAB, VF.CORS, VF.BICORSTAB, VF.CG])
def test_zero_guess_is_the_cold_solve(method):
    # a dielectric sphere: every method converges on it (CORS breaks down on
    # the plasmonic one before 1e-7)
    sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j, delta_nm=2.0, wavelength_nm=500.0)
    tol = 1e-12 if method == VF.CG else 1e-7
    with scene_solver(sc) as s:
        b = s.incident()
        x, info = s.solve(b, method, tol=tol)
        x0, info0 = s.solve_x0(b, np.zeros_like(b), method, tol=tol)
    assert info0["iterations"] == info["iterations"]
    assert np.array_equal(x0, x)


def test_solution_as_guess_takes_no_iteration():
    sc = _scene()
    with scene_solver(sc) as s:
        b = s.incident()
        x, info = s.solve(b, VF.BICGSTAB, tol=1e-8)
        x1, info1 = s.solve_x0(b, x, VF.BICGSTAB, tol=1e-8)
        assert info1["iterations"] == 0 and info1["converged"]
        assert np.array_equal(x1, x)
        # a guess that is off still reaches the tolerance on ||b - A x|| / ||b||
        rng = np.random.default_rng(2)
        g = x * (1 + 0.01 * rng.standard_normal(x.shape))
        x2, info2 = s.solve_x0(b, g, VF.BICGSTAB, tol=1e-8)
        assert info2["converged"] and info2["true_relres"] < 1e-8
        assert info2["err"] < 1e-8 and rel(x2, x) < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_neighbouring_wavelength_guess_saves_iterations(mixed):
    """A wavelength sweep started from the neighbouring wavelength's solution
    (the resident one) needs fewer iterations and reaches the same tolerance."""
    a, b = _scene(560.0), _scene(566.0)
    with scene_solver(a) as s:
        s.set_rhs_incident()
        s.solve_mixed(tol=1e-7) if mixed else s.solve_resident(VF.BICGSTAB, tol=1e-7)
        s.set_k0(b.k0)
        s.set_medium(b.material(), b.eps_table, b.flags())
        s.green_compute()
        s.set_rhs_incident()
        cold = s.solve_mixed(tol=1e-7) if mixed else s.solve_resident(VF.BICGSTAB, tol=1e-7)
        x_cold = s.solution()
        # the neighbour's solution again, as the guess
        s.set_k0(a.k0)
        s.set_medium(a.material(), a.eps_table, a.flags())
        s.green_compute()
        s.set_rhs_incident()
        s.solve_mixed(tol=1e-7) if mixed else s.solve_resident(VF.BICGSTAB, tol=1e-7)
        s.set_k0(b.k0)
        s.set_medium(b.material(), b.eps_table, b.flags())
        s.green_compute()
        s.set_rhs_incident()
        s.set_guess()
        warm = s.solve_mixed(tol=1e-7) if mixed else s.solve_resident(VF.BICGSTAB, tol=1e-7)
        x_warm = s.solution()
    assert cold["converged"] and warm["converged"]
    assert warm["true_relres"] < 1e-7
    assert warm["iterations"] < cold["iterations"], (warm, cold)
    assert rel(x_warm, x_cold) < 1e-5


def test_guess_rejected_for_bicg():
    sc = _scene()
    with scene_solver(sc) as s:
        b = s.incident()
        with pytest.raises(VF.VieError):
            s.solve_x0(b, np.zeros_like(b), VF.BICG, tol=1e-6)


def test_warm_sweep_saves_iterations_and_keeps_point_order():
    """run_sweep(warm=True) on six wavelengths of the 96^3 dimer sweep, both
    polarisations: every point converges to the same tolerance, the records
    come back in the order of the points, the cross-sections agree with the
    cold sweep's within what two solves to 1e-6 allow, and the warm sweep
    needs fewer iterations in total."""
    from vie_fft_amd import sweep as SW

    points = [(lam, psi) for lam in (560.0, 566.0, 572.0, 578.0, 584.0, 590.0) for psi in (0.0, 90.0)]
    scene = lambda lam, psi: G.dimer_case(lam, psi=psi)  # noqa: E731
    cold = SW.run_sweep(scene, points, tol=1e-6)
    warm = SW.run_sweep(scene, points, tol=1e-6, warm=True)
    assert [(r["wavelength_nm"], r["psi"]) for r in warm] == points
    assert all(r["converged"] and r["true_relres"] < 1e-6 for r in warm + cold)
    for c, w in zip(cold, warm):
        assert abs(w["q_ext"] - c["q_ext"]) <= 1e-4 * abs(c["q_ext"])
        assert abs(w["q_abs"] - c["q_abs"]) <= 1e-4 * abs(c["q_abs"])
    assert sum(r["iterations"] for r in warm) < sum(r["iterations"] for r in cold)
