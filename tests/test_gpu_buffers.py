"""Device buffers of 64 MB and more that a context frees are kept and reused
by later contexts (vie_release_buffers hands them back): a batch in a second
context, built on the first one's kept buffers, gives the same bits, and the
release reports the kept bytes once."""
import numpy as np
import pytest

from tests.gpu_util import scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def _batch_solution(sc, nb):
    with scene_solver(sc) as s:
        s.batch_reserve(nb)  # 32^3 x 3 FP64 x 64 instances = 100 MB per vector: kept on close
        for i, psi in enumerate((0.0, 90.0)):
            s.batch_set_rhs_incident(i, 0.0, 0.0, psi)
        st, infos = s.batch_solve(2, VF.BICGSTAB, tol=1e-8, mixed=True)
        assert st == VF.OK and all(i["converged"] for i in infos)
        return [s.batch_solution(i) for i in range(2)]


def test_kept_buffers_reused_and_released():
    sc = G.sphere_case(32, radius=10.0, eps_phys=-4.0 + 0.7j, delta_nm=2.0, wavelength_nm=400.0)
    VF.release_buffers()
    first = _batch_solution(sc, 64)
    second = _batch_solution(sc, 64)  # takes the first context's kept buffers
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
    released = VF.release_buffers()
    assert released >= 5 * 64 * 3 * 32 ** 3 * 16  # at least the batch's FP64 vectors
    assert VF.release_buffers() == 0
    # and a context after the release allocates afresh
    third = _batch_solution(sc, 64)
    for a, b in zip(first, third):
        assert np.array_equal(a, b)
