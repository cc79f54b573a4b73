"""Failure paths of the C ABI, all forced from the host so every kernel
exits normally.  The reference has none: a zero right-hand side makes its
BiCGSTAB compute sqrt(0/0) and loop Nx*Ny*Nz times
(vie_fft/CG_Iteration.h:734), eps = 1 gives 1/(eps-1) = inf
(vie_fft/fft.h:430), and nothing stops a diverging solve before maxit."""
import json
import os

import numpy as np
import pytest

from tests.gpu_util import scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _sphere():
    return G.sphere_case(16, radius=5.0, eps_phys=2.25 + 0.2j, delta_nm=2.0)


def _status(fn):
    try:
        fn()
    except VF.VieError as e:
        return e.status
    return VF.OK


@pytest.mark.parametrize("method", [VF.CG, VF.BICG, VF.BICGSTAB, VF.CORS, VF.BICORSTAB])
def test_zero_rhs_is_a_breakdown_status(method):
    sc = _sphere()
    with scene_solver(sc) as s:
        z = np.zeros((3,) + sc.n, np.complex128)
        assert _status(lambda: s.solve(z, method, tol=1e-6)) == VF.BREAKDOWN
        s.set_rhs(z)
        assert _status(lambda: s.solve_mixed(tol=1e-6)) == VF.BREAKDOWN


def test_maxit_returns_not_converged_with_the_iterate():
    sc = _sphere()
    with scene_solver(sc) as s:
        b = s.incident()
        x3, info = s.solve(b, VF.BICGSTAB, tol=1e-14, maxit=3)
        assert info["iterations"] == 3 and not info["converged"]
        assert np.isfinite(x3).all() and np.linalg.norm(x3) > 0
        # the same 3 iterations of a longer solve
        s.set_rhs(b)
        st = s.lib.vie_solve_resident(s.h, VF.BICGSTAB, 1e-14, 3, None)
        assert st == VF.NOT_CONVERGED
        assert np.array_equal(s.solution(), x3)


def test_non_finite_inputs_are_rejected():
    sc = _sphere()
    mat, flags = sc.material(), sc.flags()
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        bad = sc.eps_table.copy()
        bad[1] = complex(np.nan, 0.0)
        assert _status(lambda: s.set_medium(mat, bad, flags)) == VF.ERR_NAN
        e = sc.eps().copy()
        e[8, 8, 8] = complex(0.0, np.inf)
        assert _status(lambda: s.set_medium_eps(e, flags)) == VF.ERR_NAN
        s.set_medium(mat, sc.eps_table, flags)
        s.green_compute()
        b = s.incident()
        b[0, 8, 8, 8] = np.nan
        assert _status(lambda: s.set_rhs(b)) == VF.ERR_NAN
        assert _status(lambda: s.solve(b, VF.BICGSTAB, tol=1e-6)) == VF.ERR_NAN
        s.batch_reserve(2)
        assert _status(lambda: s.batch_set_rhs(1, b)) == VF.ERR_NAN


def test_eps_one_on_an_active_face_is_rejected():
    sc = _sphere()
    flags = sc.flags()
    e = sc.eps().copy()
    i, j, k = np.argwhere(flags[0] > 0)[0]
    e[i, j, k] = 1.0
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        assert _status(lambda: s.set_medium_eps(e, flags)) == VF.ERR_ARG
        assert "eps = 1" in s.lib.vie_last_error(s.h).decode()


def test_divergence_stop_ends_a_diverging_cors():
    """CORS does not converge on the 128x128x64 layered workload (its oracle
    history grows, tests/golden/cors_c2_history.json); with the divergence
    stop it ends early with DIVERGED, BiCGSTAB on the same medium converges
    unaffected, and factor 0 turns the stop off again."""
    sc = G.layered_case()
    with scene_solver(sc) as s:
        b = s.incident(*sc.rhs_angles())
        s.set_divergence_stop(1e4)
        s.set_rhs(b)
        st = s.lib.vie_solve_resident(s.h, VF.CORS, 1e-6, 0, None)
        h = s.history()
        assert st == VF.DIVERGED, VF.STATUS.get(st)
        assert len(h) < 2000 and h[-1] > 1e4 * h.min()
        info = s.solve_resident(VF.BICGSTAB, tol=1e-6, maxit=5000)
        assert info["converged"]
        s.set_divergence_stop(0)
        info = s.solve_resident(VF.CORS, tol=1e-6, maxit=len(h) + 5)
        assert info["iterations"] == len(h) + 5
        with pytest.raises(VF.VieError):
            s.set_divergence_stop(0.5)
