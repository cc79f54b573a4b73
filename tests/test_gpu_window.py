"""Slab media: the active faces fill only some planes of one axis.  The
library then reorders the axes internally (the slab's axis first) and runs
the passes of every Krylov product only over the x planes that hold active
faces; the caller's layout is unchanged.  Every result must equal the
oracle's, which works on the caller's grid with every plane:

* operator products (forward and adjoint, FP64 and FP32, host and device
  buffers) on slabs along each axis;
* 8 fixed iterations of each of the five methods (iterates to 1e-9);
* converged solves, the incident right-hand side built on the device, the
  far-field, absorbed-power and gradient outputs, and the mixed solve;
* the x pass on a shorter period than the grid's (register-resident and
  LDS-staged pencils);
* boxes limited in two or three axes (x, y and z windows, short x, y and z
  periods);
* an embedded grid (lengths without compiled kernels) with a slab;
* the same solves with VIE_WINDOW=0 (caller's axis order, every plane).

These tests pin the fixed short-period table and the pass paths it selects
(some through the VIE_*SHORT overrides), so the autotuner is off here
(VIE_TUNE=0); tuned plans are checked in test_gpu_tuning.py and run in every
other module.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import random_field, rel
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fixed_plans(monkeypatch):
    monkeypatch.setenv("VIE_TUNE", "0")

N = (16, 12, 20)
D = (0.02, 0.025, 0.03)
K0 = 2 * np.pi / 4.0
TABLE = np.array([1.1 + 0j, 2.5 - 0.3j, 4.0 - 0.8j])
SLABS = [(0, 5, 11), (1, 3, 8), (2, 6, 15)]  # axis, first plane, end plane


def slab_scene(axis, lo, hi, seed, n=N):
    """Material ids 1-2 on the planes [lo, hi) of `axis`, background 0 elsewhere;
    a face is active where both cells it joins are in the slab."""
    rng = np.random.default_rng(seed)
    mat = np.zeros(n, np.uint8)
    sl = [slice(None)] * 3
    sl[axis] = slice(lo, hi)
    mat[tuple(sl)] = rng.integers(1, 3, size=mat[tuple(sl)].shape)
    inside = mat > 0
    flags = np.zeros((3,) + tuple(n), np.uint8)
    for f in range(3):
        nb = np.zeros_like(inside)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[f] = slice(1, None)
        dst[f] = slice(0, -1)
        nb[tuple(dst)] = inside[tuple(src)]
        flags[f] = inside & nb
    return mat, TABLE[mat], flags


def box_scene(lo, hi, seed, n=N):
    """Material ids 1-2 in the box [lo, hi) (per axis), background elsewhere;
    flags as in slab_scene."""
    rng = np.random.default_rng(seed)
    mat = np.zeros(n, np.uint8)
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    mat[sl] = rng.integers(1, 3, size=mat[sl].shape)
    inside = mat > 0
    flags = np.zeros((3,) + tuple(n), np.uint8)
    for f in range(3):
        nb = np.zeros_like(inside)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[f] = slice(1, None)
        dst[f] = slice(0, -1)
        nb[tuple(dst)] = inside[tuple(src)]
        flags[f] = inside & nb
    return mat, TABLE[mat], flags


def solver(n, mat, flags, precision=VF.FP64):
    s = VF.Solver(n, D, K0, precision)
    s.set_medium(mat, TABLE, flags)
    s.green_compute()
    return s


def oracle(n, eps, flags):
    return O.Operator(n, D, K0, eps, flags, O.SCHED_FUSED)


@pytest.mark.parametrize("axis,lo,hi,order,pencil", [(0, 5, 11, (0, 1, 2), 6), (1, 3, 8, (1, 0, 2), 5),
                                                     (2, 6, 15, (2, 0, 1), 10)])
def test_plan_puts_the_slab_axis_first(axis, lo, hi, order, pencil):
    mat, eps, flags = slab_scene(axis, lo, hi, 1)
    with solver(N, mat, flags) as s:
        p = s.plan()
    assert p["axis_order"] == order and p["grid"] == tuple(N[a] for a in order)
    assert p["x_window"] == (lo, hi) and p["x_pencil"] == pencil
    os.environ["VIE_WINDOW"] = "0"
    try:
        with solver(N, mat, flags) as s:
            p = s.plan()
    finally:
        os.environ.pop("VIE_WINDOW", None)
    assert p == {"axis_order": (0, 1, 2), "x_window": (0, N[0]), "x_pencil": N[0], "grid": N, "y_window": (0, N[1]),
                 "y_pencil": N[1], "z_window": (0, N[2]), "z_pencil": N[2]}


@pytest.mark.parametrize("axis,lo,hi", SLABS)
def test_slab_operator_matches_oracle(axis, lo, hi):
    mat, eps, flags = slab_scene(axis, lo, hi, 1)
    op = oracle(N, eps, flags)
    x = random_field(N, 2) * flags
    with solver(N, mat, flags) as s:
        for adj in (False, True):
            assert rel(s.apply(x, adj), op.apply(x, adj)) <= 1e-12
    with solver(N, mat, flags, VF.FP32) as s:
        assert rel(s.apply(x), op.apply(x)) <= 3e-6


def test_slab_per_voxel_eps_matches_oracle():
    """vie_set_medium_eps (per-voxel permittivity) on a z slab: reordered axes."""
    mat, _, flags = slab_scene(2, 6, 15, 16)
    rng = np.random.default_rng(17)
    eps = rng.uniform(1.5, 6.0, N) - 1j * rng.uniform(0.0, 1.0, N)
    op = oracle(N, eps, flags)
    x = random_field(N, 18) * flags
    b = O.incident(N, D, K0, 40.0, 20.0, 10.0) * flags
    ref = op.solve(b, 3, tol=1e-300, maxit=8)
    with VF.Solver(N, D, K0) as s:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        assert s.plan()["axis_order"] == (2, 0, 1)
        assert rel(s.apply(x), op.apply(x)) <= 1e-12
        assert rel(s.apply(x, True), op.apply(x, True)) <= 1e-12
        xs, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
        assert rel(xs, ref["x"]) <= 1e-9
        s.set_rhs(b)
        s.solve_resident(VF.BICGSTAB, tol=1e-10)
        J = s.solution()
        assert abs(s.absorption() - O.absorption(N, D, K0, J, eps, flags)) <= 1e-12 * s.absorption()


@pytest.mark.parametrize("precision", [VF.FP64, VF.FP32])
def test_slab_apply_async_equals_apply(precision):
    import torch

    mat, eps, flags = slab_scene(2, 6, 15, 3)
    x = random_field(N, 4) * flags
    dt = torch.complex128 if precision == VF.FP64 else torch.complex64
    with solver(N, mat, flags, precision) as s:
        ref = s.apply(x)
        xt = torch.tensor(x, dtype=dt, device="cuda")
        yt = torch.empty_like(xt)
        s.apply_async(xt.data_ptr(), yt.data_ptr())
        torch.cuda.synchronize()
        got = yt.cpu().numpy()
    assert rel(got, ref) <= (1e-14 if precision == VF.FP64 else 1e-6)


@pytest.mark.parametrize("axis,lo,hi", SLABS)
@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_slab_fixed_iterations_match_oracle(axis, lo, hi, method):
    mat, eps, flags = slab_scene(axis, lo, hi, 5)
    b = O.incident(N, D, K0, 20.0, 30.0, 40.0) * flags
    ref = oracle(N, eps, flags).solve(b, method, tol=1e-300, maxit=8)
    with solver(N, mat, flags) as s:
        x, info = s.solve(b, method, tol=1e-300, maxit=8)
    assert info["iterations"] == ref["iterations"] == 8
    assert rel(x, ref["x"]) <= 1e-9, rel(x, ref["x"])


@pytest.mark.parametrize("axis,lo,hi", SLABS)
def test_slab_solve_and_outputs_match_oracle(axis, lo, hi):
    mat, eps, flags = slab_scene(axis, lo, hi, 6)
    op = oracle(N, eps, flags)
    ang = (20.0, 30.0, 40.0)
    b = O.incident(N, D, K0, *ang) * flags
    ref = op.solve(b, 3, tol=1e-10, maxit=2000)
    with solver(N, mat, flags) as s:
        s.set_rhs_incident(*ang)  # built on the device in the internal axis order
        info = s.solve_resident(VF.BICGSTAB, tol=1e-10)
        xs = s.solution()
        assert info["converged"] and info["true_relres"] < 1e-9
        assert abs(info["iterations"] - ref["iterations"]) <= 1
        assert rel(xs, ref["x"]) <= 1e-8
        for th, ph, ps in ((0.0, 0.0, 0.0), (25.0, 40.0, 10.0), (70.0, 200.0, 80.0)):
            fs = s.forward_scatter(th, ph, ps)
            assert abs(fs - O.forward_scatter(N, D, K0, xs, th, ph, ps)) <= 1e-11 * abs(fs)
            S = s.far_field([th], [ph])[0]
            t, p, q = np.radians([th, ph, ps])
            Z0 = 120 * np.pi
            tm = Z0 * (S[0].real * np.cos(t) * np.cos(p) + S[1].real * np.cos(t) * np.sin(p) - S[2].real * np.sin(t))
            te = Z0 * (S[0].real * np.sin(p) - S[1].real * np.cos(p))
            assert abs(np.cos(q) * tm - np.sin(q) * te - fs) <= 1e-11 * abs(fs)
        rt, rp = s.rcs(30.0)
        ot, opp = O.rcs(N, D, K0, xs, 0.0, 30.0)
        for got, want in ((rt, ot), (rp, opp)):
            ok = np.isfinite(want) & (want > np.nanmax(want) - 100)
            assert np.max(np.abs(got[ok] - want[ok])) < 1e-9
        for m in (1, 2):
            a = s.absorbed(m)
            assert abs(a - O.absorbed(N, D, K0, xs, eps, mat, m)) <= 1e-11 * abs(a)
        P = s.absorption()
        assert abs(P - O.absorption(N, D, K0, xs, eps, flags)) <= 1e-12 * abs(P)
        P2, gre, gim, ai = s.absorption_grad(tol=1e-13)
        assert ai["converged"] and P2 == P
        mixed = s.solve_mixed(tol=1e-10)
        assert mixed["converged"] and mixed["true_relres"] < 1e-10
        assert rel(s.solution(), xs) <= 1e-8
    Po, gro, gio, _ = O.absorption_grad(op, xs, tol=1e-13)
    assert gre.shape == N
    assert rel(gre, gro) <= 1e-9 and rel(gim, gio) <= 1e-9


@pytest.mark.parametrize("n,lo,hi", [((64, 8, 10), 20, 40), ((128, 6, 8), 30, 70), ((96, 8, 6), 10, 30)])
def test_short_period_register_pass_matches_oracle(n, lo, hi):
    """Windows of 20 and 40 planes in x extents 64, 96 and 128: the x pass runs
    on the register-resident 32- or 64-length pencils (window shifted to plane 0)."""
    mat, eps, flags = slab_scene(0, lo, hi, 11, n)
    b = O.incident(n, D, K0, 20.0, 30.0, 40.0) * flags
    ref = oracle(n, eps, flags).solve(b, 3, tol=1e-300, maxit=8)
    with solver(n, mat, flags) as s:
        assert s.plan()["x_pencil"] == (32 if hi - lo <= 32 else 64)
        x, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
    assert rel(x, ref["x"]) <= 1e-9, rel(x, ref["x"])
    with solver(n, mat, flags, VF.FP32) as s:
        x32, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
    assert rel(x32, ref["x"]) <= 1e-4, rel(x32, ref["x"])


# boxes limited in two or three axes: x window and y window (the BiCGSTAB
# products skip the z rows and y entries outside the latter)
BOXES = [((4, 2, 0), (12, 9, 20), (0, 1, 2), (4, 12), (2, 9)),
         ((0, 3, 5), (16, 10, 11), (2, 1, 0), (5, 11), (3, 10)),
         ((3, 0, 4), (13, 12, 18), (0, 2, 1), (3, 13), (4, 18))]


@pytest.mark.parametrize("lo,hi,order,xw,yw", BOXES)
@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_box_fixed_iterations_match_oracle(lo, hi, order, xw, yw, method):
    mat, eps, flags = box_scene(lo, hi, 14)
    b = O.incident(N, D, K0, 20.0, 30.0, 40.0) * flags
    ref = oracle(N, eps, flags).solve(b, method, tol=1e-300, maxit=8)
    with solver(N, mat, flags) as s:
        p = s.plan()
        assert p["axis_order"] == order and p["x_window"] == xw and p["y_window"] == yw
        x, info = s.solve(b, method, tol=1e-300, maxit=8)
    assert info["iterations"] == 8
    assert rel(x, ref["x"]) <= 1e-9, rel(x, ref["x"])


@pytest.mark.parametrize("lo,hi,order,xw,yw", BOXES)
def test_box_converged_and_mixed_match_oracle(lo, hi, order, xw, yw):
    mat, eps, flags = box_scene(lo, hi, 15)
    b = O.incident(N, D, K0, 10.0, 60.0, 30.0) * flags
    ref = oracle(N, eps, flags).solve(b, 3, tol=1e-10, maxit=2000)
    with solver(N, mat, flags) as s:
        s.set_rhs(b)
        info = s.solve_resident(VF.BICGSTAB, tol=1e-10)
        assert info["converged"] and abs(info["iterations"] - ref["iterations"]) <= 1
        assert rel(s.solution(), ref["x"]) <= 1e-8
        m = s.solve_mixed(tol=1e-9)
        assert m["converged"] and m["true_relres"] < 1e-9
        assert rel(s.solution(), ref["x"]) <= 1e-7


@pytest.mark.parametrize("n,lo,hi,ypen", [((16, 64, 12), (4, 20, 0), (8, 45, 12), 32),
                                          ((12, 128, 10), (3, 30, 0), (7, 90, 10), 64)])
def test_short_y_period_matches_oracle(n, lo, hi, ypen):
    """Boxes whose second-narrowest axis spans 25 of 64 and 60 of 128 planes:
    BiCGSTAB products run their y passes on register-resident 32- and
    64-length pencils (the y window shifted to entry 0)."""
    mat, eps, flags = box_scene(lo, hi, 19, n)
    b = O.incident(n, D, K0, 30.0, 40.0, 50.0) * flags
    op = oracle(n, eps, flags)
    ref = op.solve(b, 3, tol=1e-300, maxit=8)
    with solver(n, mat, flags) as s:
        p = s.plan()
        assert p["axis_order"] == (0, 1, 2) and p["y_pencil"] == ypen
        x, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
        assert rel(x, ref["x"]) <= 1e-9, rel(x, ref["x"])
        s.set_rhs(b)
        m = s.solve_mixed(tol=1e-9)
        assert m["converged"] and m["true_relres"] < 1e-9
    with solver(n, mat, flags, VF.FP32) as s:
        x32, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
    assert rel(x32, ref["x"]) <= 1e-4, rel(x32, ref["x"])


@pytest.mark.parametrize("n,lo,hi,order,zw,zpen", [((16, 12, 20), (4, 2, 5), (12, 9, 16), (0, 2, 1), (2, 9), 8),
                                                  ((12, 16, 128), (3, 2, 20), (7, 10, 100), (0, 1, 2), (20, 100), 96),
                                                  ((12, 16, 128), (3, 2, 21), (7, 10, 99), (0, 1, 2), (20, 99), 96)])
def test_short_z_period_matches_oracle(n, lo, hi, order, zw, zpen):
    """Boxes limited in all three axes: the z passes of the products inside
    solves run on the short z period (the z window, rounded down to an even
    start, shifted to entry 0), LDS-staged (8) and register-resident (96)."""
    mat, eps, flags = box_scene(lo, hi, 20, n)
    b = O.incident(n, D, K0, 35.0, 15.0, 25.0) * flags
    op = oracle(n, eps, flags)
    for method in (3, 1, 5):
        ref = op.solve(b, method, tol=1e-300, maxit=8)
        with solver(n, mat, flags) as s:
            p = s.plan()
            assert p["axis_order"] == order and p["z_window"] == zw and p["z_pencil"] == zpen
            x, _ = s.solve(b, method, tol=1e-300, maxit=8)
        assert rel(x, ref["x"]) <= 1e-9, (method, rel(x, ref["x"]))
    with solver(n, mat, flags) as s:
        s.set_rhs(b)
        m = s.solve_mixed(tol=1e-9)
        assert m["converged"] and m["true_relres"] < 1e-9
        assert rel(s.solution(), op.solve(b, 3, tol=1e-10, maxit=2000)["x"]) <= 1e-7
    with solver(n, mat, flags, VF.FP32) as s:
        x32, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
    assert rel(x32, op.solve(b, 3, tol=1e-300, maxit=8)["x"]) <= 1e-4


# Windows that end near the far edge of their axis: the window start plus the
# short period passes the grid's extent, so pencil entries at and beyond the
# last plane exist and the passes' load and store guards must drop them.
# xshort forces the x period (48: the FP32 x pass on 96-point register
# pencils, whose 24- and 36-entry variants serve windows up to 24 and 36
# planes; FP64 then takes the LDS-staged pass at 48).
EDGES = [
    # n, box lo, box hi, axis checked, xshort
    ((128, 6, 8), (84, 0, 0), (126, 6, 8), 0, None),
    ((128, 6, 8), (84, 0, 0), (126, 6, 8), 0, "48"),
    ((64, 8, 10), (40, 0, 0), (62, 8, 10), 0, "48"),
    ((64, 8, 10), (30, 0, 0), (62, 8, 10), 0, "48"),
    ((96, 8, 6), (50, 0, 0), (96, 8, 6), 0, "48"),
    ((12, 128, 10), (3, 70, 0), (7, 126, 10), 1, None),
    ((12, 16, 128), (3, 2, 60), (7, 10, 126), 2, None),
]


@pytest.mark.parametrize("n,lo,hi,axis,xshort", EDGES)
def test_short_period_past_the_far_edge_matches_oracle(monkeypatch, n, lo, hi, axis, xshort):
    if xshort:
        monkeypatch.setenv("VIE_XSHORT", xshort)
    mat, eps, flags = box_scene(lo, hi, 23, n)
    b = O.incident(n, D, K0, 25.0, 35.0, 45.0) * flags
    ref = oracle(n, eps, flags).solve(b, 3, tol=1e-300, maxit=8)
    key = ("x", "y", "z")[axis]
    for prec, bound in ((VF.FP64, 1e-9), (VF.FP32, 1e-4)):
        with solver(n, mat, flags, prec) as s:
            p = s.plan()
            assert p["axis_order"] == (0, 1, 2), p
            w0, w1 = p[key + "_window"]
            pen = p[key + "_pencil"]
            assert pen < n[axis] and w0 + pen > n[axis], p
            if xshort:
                assert pen == int(xshort)
            x, info = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
        assert info["iterations"] == 8
        assert rel(x, ref["x"]) <= bound, (prec, rel(x, ref["x"]))


def test_embedded_slab_matches_oracle():
    n = (18, 14, 11)  # compiled as 20 x 16 x 12
    assert not VF.supported_n(18) and not VF.supported_n(14) and not VF.supported_n(11)
    mat, eps, flags = slab_scene(2, 3, 8, 7, n)
    op = oracle(n, eps, flags)
    x = random_field(n, 8) * flags
    b = O.incident(n, D, K0) * flags
    ref = op.solve(b, 3, tol=1e-10, maxit=2000)
    with solver(n, mat, flags) as s:
        assert rel(s.apply(x), op.apply(x)) <= 1e-12
        assert rel(s.apply(x, True), op.apply(x, True)) <= 1e-12
        xs, info = s.solve(b, VF.BICGSTAB, tol=1e-10)
        assert abs(info["iterations"] - ref["iterations"]) <= 1
        assert rel(xs, ref["x"]) <= 1e-8
        assert abs(s.absorbed(1) - O.absorbed(n, D, K0, xs, eps, mat, 1)) <= 1e-11 * abs(s.absorbed(1))


def test_one_context_follows_a_new_medium_and_k0():
    """One context: a z slab (internal order z, x, y) solved in mixed precision,
    then an x slab (caller's order) after G~ was computed -- the operators,
    G~ and the short period are rebuilt for the new order without a second
    green_compute -- then a new wavelength."""
    mat_z, eps_z, fl_z = slab_scene(2, 6, 15, 12)
    mat_x, eps_x, fl_x = slab_scene(0, 5, 11, 13)
    b_z = O.incident(N, D, K0) * fl_z
    b_x = O.incident(N, D, K0, 30.0, 10.0, 20.0) * fl_x
    with solver(N, mat_z, fl_z) as s:
        s.set_rhs(b_z)
        m = s.solve_mixed(tol=1e-9)
        assert m["converged"] and s.plan()["axis_order"] == (2, 0, 1)
        assert rel(s.solution(), oracle(N, eps_z, fl_z).solve(b_z, 3, tol=1e-10, maxit=2000)["x"]) <= 1e-7
        s.set_medium(mat_x, TABLE, fl_x)
        assert s.plan()["axis_order"] == (0, 1, 2) and s.plan()["x_window"] == (5, 11)
        x, info = s.solve(b_x, VF.BICGSTAB, tol=1e-300, maxit=8)
        ref = oracle(N, eps_x, fl_x).solve(b_x, 3, tol=1e-300, maxit=8)
        assert rel(x, ref["x"]) <= 1e-9
        s.set_rhs(b_x)
        m = s.solve_mixed(tol=1e-9)
        assert m["converged"]
        k1 = 2 * np.pi / 3.3
        s.set_k0(k1)
        s.green_compute()
        x, _ = s.solve(b_x, VF.BICGSTAB, tol=1e-300, maxit=8)
        ref = O.Operator(N, D, k1, eps_x, fl_x, O.SCHED_FUSED).solve(b_x, 3, tol=1e-300, maxit=8)
        assert rel(x, ref["x"]) <= 1e-9


@pytest.mark.parametrize("axis,lo,hi", SLABS)
def test_window_off_gives_the_same_solves(axis, lo, hi):
    mat, eps, flags = slab_scene(axis, lo, hi, 9)
    b = O.incident(N, D, K0, 10.0, 50.0, 0.0) * flags
    out = {}
    for w in ("1", "0"):
        os.environ["VIE_WINDOW"] = w
        try:
            with solver(N, mat, flags) as s:
                x8, _ = s.solve(b, VF.BICGSTAB, tol=1e-300, maxit=8)
                s.set_rhs(b)
                mixed = s.solve_mixed(tol=1e-9)
                out[w] = (x8, s.solution(), mixed)
        finally:
            os.environ.pop("VIE_WINDOW", None)
    assert rel(out["1"][0], out["0"][0]) <= 1e-11
    assert out["1"][2]["converged"] and out["0"][2]["converged"]
    assert rel(out["1"][1], out["0"][1]) <= 1e-8
