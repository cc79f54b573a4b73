"""Per-instance operators of a batch when the tuner gave the FP64 and FP32
operators different short periods.  The batch set-up computes every
instance's Green's-function octants on each operator's own window grid (they
were once computed on the FP32 grid only and copied into both, which read
past the end of that buffer for a longer FP64 period).  The plans are pinned
through a plan cache that this test writes: the scene is tuned in one child
process, the cache is rewritten with the full x period for FP64 and a short
one for FP32, and a second child solves three instances as one mixed batch
and one after the other (tests/batch_plans_child.py).

Also: per-instance operators set again after the medium went to another
geometry and back, when the batch buffers are kept but the short-period
tables they point into were rebuilt (the batch products once kept the old,
freed addresses)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHILD = os.path.join(ROOT, "tests", "batch_plans_child.py")


def _child(stage, cache):
    env = dict(os.environ, VIE_TUNE_CACHE=str(cache))
    print("child stage", stage, flush=True)
    r = subprocess.run([sys.executable, "-u", CHILD, stage], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    return json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])


def test_batch_with_different_fp64_and_fp32_periods(tmp_path):
    cache = tmp_path / "plans.txt"
    _child("tune", cache)
    lines = [ln.split() for ln in cache.read_text().splitlines() if ln.strip()]
    keys = {ln[0]: ln for ln in lines}
    assert set(keys) == {"f64", "f32"}, lines
    out = []
    for prec, ln in keys.items():
        n0, wx0, wx1 = int(ln[1]), int(ln[4]), int(ln[5])
        var, ns, nys, nzs, us, usd = ln[11:17]
        if prec == "f64":
            ns = 0  # the full x period
        else:
            ns = next(m for m in range(wx1 - wx0, n0) if VF.supported_n(m))  # the shortest one
        out.append(" ".join(ln[:10]) + f" | {var} {ns} {nys} {nzs} {us} {usd}")
    cache.write_text("\n".join(out) + "\n")
    r = _child("check", cache)
    assert r["tuning"]["fp64"]["pencils"][0] != r["tuning"]["fp32"]["pencils"][0], r["tuning"]
    assert r["status"] == VF.OK and all(r["converged"]), r
    assert max(r["rel"]) <= 1e-8, r["rel"]


def _solve_batch(s, scenes):
    s.batch_set_operators(0, [q.k0 for q in scenes], np.stack([q.eps_table for q in scenes]))
    for i in range(len(scenes)):
        s.batch_set_rhs_incident(i)
    st, infos = s.batch_solve(len(scenes), VF.BICGSTAB, tol=1e-9, mixed=True)
    assert st == VF.OK and all(i["converged"] for i in infos), (st, infos)
    return [s.batch_solution(i) for i in range(len(scenes))]


def test_batch_after_a_medium_round_trip(monkeypatch):
    """Per-instance operators set again after the medium went to another
    geometry and back: the short-period tables were rebuilt at new addresses
    in between (the two geometries' default periods differ) while the batch
    buffers, sized for the same lengths, were kept.  The second batch must
    read the rebuilt tables and give the first batch's solutions exactly."""
    monkeypatch.setenv("VIE_TUNE", "0")  # each geometry's default periods, no timing
    scenes = [G.sphere_case(32, radius=5.0, eps_phys=e, delta_nm=4.0, wavelength_nm=w)
              for e, w in ((4.0 + 0.5j, 400.0), (3.0 + 0.2j, 450.0))]
    big = G.sphere_case(32, radius=12.0, eps_phys=4.0 + 0.5j, delta_nm=4.0, wavelength_nm=400.0)
    a = scenes[0]
    s = VF.Solver(a.n, a.d, a.k0, VF.FP64)
    try:
        s.set_medium(a.material(), a.eps_table, a.flags())
        s.green_compute()
        s.batch_reserve(2)
        first = _solve_batch(s, scenes)
        plan_a = s.plan()
        s.set_medium(big.material(), big.eps_table, big.flags())
        s.green_compute()
        plan_big = s.plan()
        s.set_medium(a.material(), a.eps_table, a.flags())
        s.green_compute()
        assert s.plan() == plan_a
        second = _solve_batch(s, scenes)
    finally:
        s.close()
    assert plan_big["x_pencil"] != plan_a["x_pencil"], (plan_a, plan_big)
    for x, y in zip(first, second):
        assert np.array_equal(x, y), float(np.abs(x - y).max())
