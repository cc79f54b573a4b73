"""Per-launch kernel timing (vie_time_kernel / vie_kernel_time): every kernel
class is sampled during a solve, the times are positive and below the solve's
own per-iteration time, and timing never changes a result (the sampled
launches record their events with the dispatch itself)."""
import numpy as np
import pytest

from tests.gpu_util import scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def test_every_class_is_timed_without_changing_results():
    sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j)
    with scene_solver(sc) as s:
        s.set_rhs_incident()
        ref64 = s.solve_resident(VF.BICGSTAB, tol=1e-8)
        x64 = s.solution()
        refm = s.solve_mixed(tol=1e-8)
        xm = s.solution()
        for cls in (1, 2, 3, 4, 5, 6, 7):
            s.time_kernel(cls)
            info = s.solve_resident(VF.BICGSTAB, tol=1e-8)
            n, ms = s.kernel_time()
            s.time_kernel(0)
            assert info["iterations"] == ref64["iterations"] and np.array_equal(s.solution(), x64)
            assert n >= 1 and 0 < ms < 1e3 * info["seconds"] / info["iterations"], (cls, n, ms)
        for cls in (11, 12, 13, 14, 15, 16, 17):
            s.time_kernel(cls)
            info = s.solve_mixed(tol=1e-8)
            n, ms = s.kernel_time()
            s.time_kernel(0)
            assert info["iterations"] == refm["iterations"] and np.array_equal(s.solution(), xm)
            assert n >= 1 and 0 < ms < 1e3 * info["seconds"] / info["iterations"], (cls, n, ms)
