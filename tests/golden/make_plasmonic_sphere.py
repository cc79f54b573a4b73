"""Writes plasmonic_sphere_q.json: the CPU oracle's Q_ext and Q_abs of the
plasmonic sphere workload (radius 20 nm, eps = -4 + 0.7i physics form,
400 nm, x-polarised along +z) at 10, 15 and 20 cells per radius, solved by
BiCGSTAB to 1e-8, beside the Mie series (Mie_LPR_Multilayer.m:138-177
restated in oracle/mie.py).  The GPU test compares its own cross-sections
with these; the oracle solves take a few minutes on one core.

    python tests/golden/make_plasmonic_sphere.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import mie  # noqa: E402
from oracle import oracle as O  # noqa: E402
from vie_fft_amd import geometry as G  # noqa: E402

Z0 = 120 * np.pi
CASES = [(32, 10.0), (48, 15.0), (64, 20.0)]


def main():
    qe_m, _, qa_m = mie.efficiencies(20, 400, -4 + 0.7j)
    out = {"radius_nm": 20.0, "wavelength_nm": 400.0, "eps_phys": [-4.0, 0.7], "tol": 1e-8,
           "mie": {"q_ext": qe_m, "q_abs": qa_m}, "cases": []}
    for n, R in CASES:
        sc = G.sphere_case(n, radius=R, eps_phys=-4 + 0.7j, delta_nm=20.0 / R)
        mat, fl, eps = sc.material(), sc.flags(), sc.eps()
        op = O.Operator(sc.n, sc.d, sc.k0, eps, fl, O.SCHED_FUSED)
        b = O.incident(sc.n, sc.d, sc.k0) * fl
        res = op.solve(b, 3, tol=1e-8)
        J = res["x"]
        qe = O.forward_scatter(sc.n, sc.d, sc.k0, J) / sc.area
        qa = O.absorbed(sc.n, sc.d, sc.k0, J, eps, mat, 1) * Z0 / sc.area
        out["cases"].append({"n": n, "radius_cells": R, "iterations": res["iterations"], "q_ext": qe, "q_abs": qa})
        print(out["cases"][-1], flush=True)
    with open(os.path.join(HERE, "plasmonic_sphere_q.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
