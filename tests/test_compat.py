"""The reference-API compat header (include/vie_compat.hpp): a driver that
uses the reference's globals and call sequence builds against it, and on the
GPU gives exactly what the C ABI gives through the Python front end."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.compat import build_driver as BD
from vie_fft_amd import geometry as G

LIB = os.path.join(BD.ROOT, "vie_fft_amd", "libviefft.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libviefft.so not built")
def test_compat_driver_builds(tmp_path):
    BD.build(str(tmp_path / "driver.bin"))
    assert os.path.getsize(tmp_path / "driver.bin") > 0


def _write_case(path, sc, eps_table, material, flags, angles):
    with open(path, "wb") as f:
        f.write(np.array(sc.n, np.int32).tobytes())
        f.write(np.array(list(sc.d) + [sc.k0] + list(angles), np.float64).tobytes())
        f.write(np.array([len(eps_table)], np.int32).tobytes())
        f.write(np.ascontiguousarray(eps_table, np.complex128).tobytes())
        f.write(np.ascontiguousarray(material, np.uint8).tobytes())
        f.write(np.ascontiguousarray(flags, np.uint8).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "layered"])
def test_compat_driver_matches_c_abi(tmp_path, case):
    """The layered medium (bands along z) runs on the library's reordered
    internal axes; the sphere keeps the caller's order."""
    from vie_fft_amd import viefft as VF

    drv = BD.OUT if os.path.exists(BD.OUT) else BD.build(str(tmp_path / "driver.bin"))
    if case == "sphere":
        sc = G.sphere_case(n=20, radius=6.0, eps_phys=-4.0 + 0.7j, wavelength_nm=450.0)
    else:
        sc = G.layered_case(n=(24, 20, 32), band_start=4, sphere_z=14.0)
    mat, flags = sc.material(), sc.flags()
    angles = (0.0, 0.0, 30.0)
    case, out = tmp_path / "case.bin", tmp_path / "out.bin"
    _write_case(case, sc, sc.eps_table, mat, flags, angles)
    r = subprocess.run([drv, str(case), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps = int(re.search(r"Steps: (\d+)", r.stdout).group(1))
    J = np.fromfile(out, np.complex128).reshape((3,) + sc.n)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(mat, sc.eps_table, flags)
        s.green_compute()
        x, info = s.solve(s.incident(*angles), method=VF.BICGSTAB, tol=1e-3)
    assert steps == info["iterations"] - 1
    assert np.array_equal(J, x)


@pytest.mark.gpu
def test_compat_operator_calls_match_c_abi_and_oracle(tmp_path):
    """X_FFT/Y_FFT/Z_FFT, Current_Reset and the _Conjugate set with the
    reference's signatures (vie_fft/CG_Iteration.h:86-173, fft.h:13-35) on an
    unmasked field: equal to vie_apply_raw / vie_apply, and to the oracle's
    product (all flags set, i.e. no Current_Reset, for the raw ones)."""
    from oracle import oracle as O
    from vie_fft_amd import viefft as VF

    drv = BD.OUT if os.path.exists(BD.OUT) else BD.build(str(tmp_path / "driver.bin"))
    sc = G.layered_case(n=(24, 20, 32), band_start=4, sphere_z=14.0)
    mat, flags, eps = sc.material(), sc.flags(), sc.eps()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)
    case, xb, out = tmp_path / "case.bin", tmp_path / "x.bin", tmp_path / "out.bin"
    _write_case(case, sc, sc.eps_table, mat, flags, (0.0, 0.0, 0.0))
    x.astype(np.complex128).tofile(xb)
    r = subprocess.run([drv, "op", str(case), str(xb), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw, masked, conj = np.fromfile(out, np.complex128).reshape((3, 3) + sc.n)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(mat, sc.eps_table, flags)
        s.green_compute()
        assert np.array_equal(raw, s.apply_raw(x))
        assert np.array_equal(masked, s.apply(x))
        assert np.array_equal(conj, s.apply_raw(x, adjoint=True))
    ones = np.ones_like(flags)
    op_all = O.Operator(sc.n, sc.d, sc.k0, eps, ones, O.SCHED_FUSED)
    op = O.Operator(sc.n, sc.d, sc.k0, eps, flags, O.SCHED_FUSED)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    assert rel(raw, op_all.apply(x)) <= 1e-12
    assert rel(masked, op.apply(x)) <= 1e-12
    assert rel(conj, op_all.apply(x, adjoint=True)) <= 1e-12
