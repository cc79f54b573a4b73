"""Writes cors_c2_history.json: the CPU oracle's CORS residual history
(||r_k|| / ||b|| after each of the first 60 iterations, the stop quantity of
vie_fft/CG_Iteration.h:1040-1050) on the 128x128x64 layered workload (five
materials, 550 nm), where CORS does not converge.  The GPU test compares
its own history with it.  About ten minutes on one core.

    python tests/golden/make_cors_history.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from vie_fft_amd import geometry as G  # noqa: E402

ITERS = 60


def main():
    sc = G.layered_case()
    eps, flags = sc.eps(), sc.flags()
    op = O.Operator(sc.n, sc.d, sc.k0, eps, flags, O.SCHED_FUSED)
    b = O.incident(sc.n, sc.d, sc.k0) * flags
    out = op.solve(b, 4, tol=1e-300, maxit=ITERS, history=True)
    hist = [float(v) for v in out["history"]]
    with open(os.path.join(HERE, "cors_c2_history.json"), "w") as f:
        json.dump({"workload": "128x128x64 layered slab + 4x4 spheres, 550 nm, x-pol", "method": "CORS",
                   "iterations": out["iterations"], "history": hist}, f, indent=1)
    print(hist)


if __name__ == "__main__":
    main()
