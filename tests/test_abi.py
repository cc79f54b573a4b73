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
