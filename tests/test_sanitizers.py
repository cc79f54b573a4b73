"""The oracle's CPU tests under AddressSanitizer and UndefinedBehaviorSanitizer:
the oracle is rebuilt with -fsanitize=address,undefined (oracle/Makefile,
target sanitize) and the tests run in a child python with libasan preloaded;
any sanitizer report aborts the child (-fno-sanitize-recover=all)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "oracle", "_san", "libvie_oracle.so")


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_tests_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    ubsan = _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["VIE_ORACLE_LIB"] = SAN_LIB
    env["LD_PRELOAD"] = f"{asan}:{ubsan}"
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    tests = ["tests/test_oracle_fft.py", "tests/test_oracle_green.py", "tests/test_oracle_operator.py",
             "tests/test_oracle_krylov.py", "tests/test_oracle_adjoint.py", "tests/test_oracle_farfield.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests], cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
