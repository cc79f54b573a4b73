"""Child process of tests/test_gpu_batch_plans.py (a plan cache is read once
per process, so each stage runs in a fresh one with VIE_TUNE_CACHE set).

stage "tune": tune the scene's FP64 and FP32 operators into the cache file.
stage "check": with the cache the parent rewrote (different x periods for the
two precisions), solve three instances (own wavenumber and permittivity
table each) as one mixed-precision batch and one after the other in fresh
contexts; print both results as one JSON line."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from vie_fft_amd import geometry as G  # noqa: E402
from vie_fft_amd import viefft as VF  # noqa: E402

WAVELENGTHS = (400.0, 450.0, 500.0)
EPS = (4.0 + 0.5j, 3.0 + 0.2j, 5.0 + 1.0j)


def scene(i=0):
    return G.sphere_case(32, radius=5.0, eps_phys=EPS[i], delta_nm=4.0, wavelength_nm=WAVELENGTHS[i])


def context(sc):
    s = VF.Solver(sc.n, sc.d, sc.k0, VF.FP64)
    s.set_medium(sc.material(), sc.eps_table, sc.flags())
    s.green_compute()
    return s


def main(stage):
    sc = scene()
    if stage == "tune":
        with context(sc) as s:
            s.set_rhs_incident()
            s.solve_mixed(tol=1e-6)  # makes (and tunes) the FP32 operator too
            print(json.dumps({"tuning": s.tuning(), "plan": s.plan()}))
        return
    scenes = [scene(i) for i in range(3)]
    with context(sc) as s:
        s.batch_reserve(3)
        s.batch_set_operators(0, [q.k0 for q in scenes], np.stack([q.eps_table for q in scenes]))
        for i in range(3):
            s.batch_set_rhs_incident(i)
        st, infos = s.batch_solve(3, VF.BICGSTAB, tol=1e-10, mixed=True)
        batched = [s.batch_solution(i) for i in range(3)]
        tuning = s.tuning()
    single = []
    for q in scenes:
        with context(q) as s:
            s.set_rhs_incident()
            s.solve_mixed(tol=1e-10)
            single.append(s.solution())
    rel = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(batched, single)]
    print(json.dumps({"tuning": tuning, "status": st, "converged": [bool(i["converged"]) for i in infos],
                      "rel": rel}))


if __name__ == "__main__":
    main(sys.argv[1])
