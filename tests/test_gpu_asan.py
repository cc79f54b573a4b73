"""The library's host code under AddressSanitizer: tests/asan/abi_driver.bin
(built by tests/asan/build.py, part of __graft_entry__.build()) calls the
whole C ABI -- set-up, products, the five methods, mixed, warm-started,
preconditioned and batched solves, every output, the adjoint gradient, an
embedded grid and the argument checks -- against a library whose host code
is instrumented; any sanitizer report fails the run (exit code 23, no abort)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "asan", "abi_driver.bin")


def test_abi_under_address_sanitizer():
    assert os.path.exists(DRIVER), "tests/asan/abi_driver.bin is missing: run tests/asan/build.py"
    env = dict(os.environ)
    # leaks: the HIP runtime's own allocations live to exit; container
    # annotations: std::vector is shared with uninstrumented translation units
    env["ASAN_OPTIONS"] = ("detect_leaks=0:detect_container_overflow=0:abort_on_error=0:halt_on_error=1:"
                           "exitcode=23:verify_asan_link_order=0")
    print("running", DRIVER, flush=True)
    r = subprocess.run([DRIVER], env=env, capture_output=True, text=True, timeout=240)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "abi_driver ok" in r.stdout, out[-4000:]
