"""The C-ABI library loads on the CPU and exports every symbol include/viefft.h declares."""
import os
import re
import subprocess

import pytest

from vie_fft_amd import viefft as VF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "viefft.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char\*)\s+(vie_\w+)\s*\(", text, re.M)))


def test_header_and_wrapper_agree():
    assert _declared() == sorted(VF.EXPORTS)


def test_library_exports_every_symbol():
    lib = VF.load()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", VF.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (vie_\w+)", out))
    assert set(_declared()) <= exported


def test_size_table_without_gpu():
    assert VF.supported_n(128) and VF.supported_n(96) and VF.supported_n(256)
    assert not VF.supported_n(7) and not VF.supported_n(512)


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(VF.VieError):
        VF.Solver((8, 8, 8), (0.02,) * 3, 1.0)


NAN = float("nan")


@pytest.mark.parametrize("n,d,k0,precision,status", [
    ((257, 8, 8), (0.02,) * 3, 1.0, VF.FP64, VF.ERR_UNSUPPORTED),   # beyond the largest compiled length
    ((8, 8, 1000), (0.02,) * 3, 1.0, VF.FP32, VF.ERR_UNSUPPORTED),
    ((0, 8, 8), (0.02,) * 3, 1.0, VF.FP64, VF.ERR_ARG),             # empty grid
    ((8, 8, 8), (0.02, 0.0, 0.02), 1.0, VF.FP64, VF.ERR_ARG),       # zero cell
    ((8, 8, 8), (0.02, NAN, 0.02), 1.0, VF.FP64, VF.ERR_ARG),
    ((8, 8, 8), (0.02,) * 3, 0.0, VF.FP64, VF.ERR_ARG),             # zero wavenumber
    ((8, 8, 8), (0.02,) * 3, float("inf"), VF.FP64, VF.ERR_ARG),
    ((8, 8, 8), (0.02,) * 3, 1.0, 7, VF.ERR_ARG),                   # unknown precision
])
def test_create_rejects_bad_descriptors_before_any_device_call(n, d, k0, precision, status):
    # vie_create checks its descriptor before it touches a device, so these
    # statuses come back the same with or without a GPU
    with pytest.raises(VF.VieError) as e:
        VF.Solver(n, d, k0, precision)
    assert e.value.status == status
