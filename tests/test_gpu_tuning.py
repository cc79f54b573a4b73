"""The autotuner: per geometry and precision it times the pass variants and
short periods of the solvers' windowed product on the device and keeps the
fastest; the fixed defaults are its starting point, so its plan is never
slower than them (beyond timing noise), and the operator stays the
reference's (parity against the oracle holds for whatever plan it picks)."""
import os

import numpy as np
import pytest

from tests.gpu_util import oracle_op, rel, scene_solver
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["layered", "dimer"])
def test_tuned_plan_is_not_slower_and_exact(case):
    sc = G.layered_case(n=(64, 64, 32), band_start=4, sphere_z=20.0) if case == "layered" else G.dimer_case(600.0)
    with scene_solver(sc) as s:
        s.set_rhs_incident()
        s.solve_mixed(tol=1e-3, maxit=4)  # creates the FP32 operator (tuned as it is made)
        t = s.tuning()
        print(case, t, flush=True)
        for name in ("fp64", "fp32"):
            p = t[name]
            assert p["tuned"] and 0 <= p["variant"] < 8
            assert p["us_per_product"] <= 1.05 * p["us_per_product_defaults"], p
        if case == "layered":
            eps, flags = sc.eps(), sc.flags()
            rng = np.random.default_rng(1)
            x = (rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)) * flags
            assert rel(s.apply(x), oracle_op(sc.n, sc.d, sc.k0, eps, flags).apply(x)) <= 1e-12
            b = s.incident() * flags
            xs, info = s.solve(b, VF.BICGSTAB, tol=1e-30, maxit=8)
            ref = oracle_op(sc.n, sc.d, sc.k0, eps, flags).solve(b, 3, tol=1e-30, maxit=8)
            assert rel(xs, ref["x"]) <= 1e-9


def test_plans_are_kept_per_geometry_and_can_be_turned_off(monkeypatch):
    sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j)
    with scene_solver(sc) as a:
        pa = a.tuning()["fp64"]
    with scene_solver(sc) as b:
        pb = b.tuning()["fp64"]
    assert pa == pb  # the second context reuses the first one's plan
    monkeypatch.setenv("VIE_TUNE", "0")
    with scene_solver(sc) as c:
        pc = c.tuning()["fp64"]
    assert not pc["tuned"] and pc["variant"] == -1
