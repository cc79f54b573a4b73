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


_CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from tests.gpu_util import oracle_op, rel, scene_solver
from vie_fft_amd import geometry as G
sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j)
with scene_solver(sc) as s:
    t = s.tuning()
    rng = np.random.default_rng(3)
    flags = sc.flags()
    x = (rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)) * flags
    err = rel(s.apply(x), oracle_op(sc.n, sc.d, sc.k0, sc.eps(), flags).apply(x))
print(json.dumps({"tuning": t, "err": err}))
"""


def test_damaged_cache_line_is_tuned_again(tmp_path, monkeypatch):
    """A VIE_TUNE_CACHE line whose short period is shorter than the window
    (it would change the operator) is ignored: the context tunes again and
    its products equal the oracle's.  The cache file is read once per
    process, so the check runs in a child process."""
    import json
    import subprocess
    import sys

    sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j)
    monkeypatch.setenv("VIE_TUNE", "0")
    with scene_solver(sc) as s:
        p = s.plan()
    monkeypatch.delenv("VIE_TUNE")
    w = p["x_window"][1] - p["x_window"][0]
    bad = max(m for m in range(2, w) if VF.supported_n(m))
    key = " ".join(str(v) for v in (*p["grid"], *p["x_window"], *p["y_window"], *p["z_window"]))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

    def child(line):
        cache = tmp_path / "tune.cache"
        cache.write_text("".join(f"{pr} {key} | {line}\n" for pr in ("f64", "f32")))
        env = dict(os.environ, VIE_TUNE_CACHE=str(cache))
        r = subprocess.run([sys.executable, "-c", _CHILD, root], env=env, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        out = json.loads(r.stdout.strip().splitlines()[-1])
        assert out["err"] <= 1e-12
        return out["tuning"]["fp64"]

    # a valid line is taken as it stands (so the key matches this geometry) ...
    t = child("5 0 0 0 1.000 2.000")
    assert t["tuned"] and t["variant"] == 5 and t["us_per_product"] == 1.0, t
    # ... a line with a period shorter than the window is not
    t = child(f"0 {bad} 0 0 1.000 2.000")
    assert t["tuned"] and t["pencils"][0] >= w > bad and t["us_per_product"] != 1.0, t
