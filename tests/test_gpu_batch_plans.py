"""Per-instance operators of a batch when the tuner gave the FP64 and FP32
operators different short periods.  The batch set-up computes every
instance's Green's-function octants on each operator's own window grid (they
were once computed on the FP32 grid only and copied into both, which read
past the end of that buffer for a longer FP64 period).  The plans are pinned
through a plan cache that this test writes: the scene is tuned in one child
process, the cache is rewritten with the full x period for FP64 and a short
one for FP32, and a second child solves three instances as one mixed batch
and one after the other (tests/batch_plans_child.py)."""
import json
import os
import subprocess
import sys

import pytest

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
