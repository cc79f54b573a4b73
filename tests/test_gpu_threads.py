"""One context per host thread: contexts used from two threads at once (each
call releases the interpreter lock, so the library is entered concurrently)
give the bits the same solves give one after the other.  The sequential runs
come first, so both threads find their plans in the process's tuning cache."""
import threading

import numpy as np
import pytest

from tests.gpu_util import scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def _sphere_fp64():
    sc = G.sphere_case(24, radius=8.0, eps_phys=4.0 + 0.5j, delta_nm=4.0, wavelength_nm=500.0)
    with scene_solver(sc) as s:
        s.set_rhs_incident(*sc.rhs_angles())
        info = s.solve_resident(VF.BICGSTAB, tol=1e-10)
        return s.solution(), info["iterations"], s.forward_scatter(*sc.rhs_angles())


def _layered_mixed():
    sc = G.layered_case(n=(32, 32, 52), band_start=2, sphere_z=30.0)
    with scene_solver(sc) as s:
        s.set_rhs_incident(*sc.rhs_angles())
        info = s.solve_mixed(tol=1e-8)
        return s.solution(), info["iterations"], s.absorbed(sc.absorb_mat)


def test_two_threads_equal_sequential():
    ref = [_sphere_fp64(), _layered_mixed()]
    got = [[], []]
    errors = []

    def run(k, fn):
        try:
            for _ in range(3):
                got[k].append(fn())
        except Exception as e:  # the assertion below reports it
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(0, _sphere_fp64)),
               threading.Thread(target=run, args=(1, _layered_mixed))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for k in range(2):
        assert len(got[k]) == 3
        for x, its, q in got[k]:
            assert its == ref[k][1]
            assert np.array_equal(x, ref[k][0])
            assert q == ref[k][2]
