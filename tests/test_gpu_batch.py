"""Batched solves (many instances of one geometry in one set of launches)
against one-at-a-time solves of the same instances.

The reference solves one wavelength per process (vie_fft/main.cpp:27-36,
vie_fft/CG_FFT_VIE.cpp:45-61); the batch must give every instance exactly
what its own solve gives: each instance runs the same kernels on the same
data with the same reduction order, so iterates, iteration counts and
cross-sections are compared at roundoff level (FP64 bit patterns are
expected to agree; the bound leaves room for nothing more than that).
"""
import numpy as np
import pytest

from vie_fft_amd import geometry as G
from vie_fft_amd import sweep as SW
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu

LAMS = (450.0, 560.0, 700.0)
PSIS = (0.0, 90.0)


def drude_sphere(lam, psi=0.0, n=24, radius=7.0):
    sc = G.sphere_case(n, radius=radius, eps_phys=complex(G.drude_eps_phys(lam)), delta_nm=2.0,
                       wavelength_nm=lam, psi=psi)
    return sc


def single(sc, psi, mixed, method=VF.BICGSTAB, tol=1e-8, maxit=0, precision=VF.FP64):
    with VF.Solver(sc.n, sc.d, sc.k0, precision) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.set_rhs_incident(sc.theta, sc.phi, psi)
        info = s.solve_mixed(tol=tol, maxit=maxit) if mixed else s.solve_resident(method, tol=tol, maxit=maxit)
        x = s.solution()
        q = s.forward_scatter(sc.theta, sc.phi, psi)
        a = s.absorbed(1)
    return x, info, q, a


def batch_of(points, mixed, method=VF.BICGSTAB, tol=1e-8, maxits=None, precision=VF.FP64, tols=None):
    scenes = [drude_sphere(lam, psi) for lam, psi in points]
    sc0 = scenes[0]
    with VF.Solver(sc0.n, sc0.d, sc0.k0, precision) as s:
        s.set_medium(sc0.material(), sc0.eps_table, sc0.flags())
        s.green_compute()
        s.batch_reserve(len(points))
        s.batch_set_operators(0, [sc.k0 for sc in scenes], np.stack([sc.eps_table for sc in scenes]))
        for i, (sc, (lam, psi)) in enumerate(zip(scenes, points)):
            s.batch_set_rhs_incident(i, sc.theta, sc.phi, psi)
        st, infos = s.batch_solve(len(points), method=method, tol=tol, mixed=mixed, maxits=maxits, tols=tols)
        xs = [s.batch_solution(i) for i in range(len(points))]
        qs = [s.batch_forward_scatter(i, sc.theta, sc.phi, psi) for i, (sc, (lam, psi)) in
              enumerate(zip(scenes, points))]
        ab = [s.batch_absorbed(i, 1) for i in range(len(points))]
        # all instances' outputs with one read-back: the same values
        fw, ab2 = s.batch_outputs([sc.theta for sc in scenes], [sc.phi for sc in scenes], [p for _, p in points], 1)
        assert list(fw) == qs and list(ab2) == ab
        fw1, none = s.batch_outputs([sc.theta for sc in scenes], [sc.phi for sc in scenes], [p for _, p in points])
        assert list(fw1) == qs and none is None
    return st, infos, xs, qs, ab, scenes


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_wavelengths_equal_single_solves(mixed):
    points = [(lam, psi) for lam in LAMS for psi in PSIS]
    print(f"batch of {len(points)} instances, mixed={mixed}", flush=True)
    st, infos, xs, qs, ab, scenes = batch_of(points, mixed)
    assert st == VF.OK
    for i, (lam, psi) in enumerate(points):
        x, info, q, a = single(scenes[i], psi, mixed)
        assert infos[i]["converged"] and info["converged"]
        assert infos[i]["iterations"] == info["iterations"], (i, infos[i], info)
        assert infos[i]["outer"] == info["outer"]
        assert rel(xs[i], x) <= 1e-13, (i, rel(xs[i], x))
        assert abs(qs[i] - q) <= 1e-12 * abs(q)
        assert abs(ab[i] - a) <= 1e-12 * abs(a)
        assert infos[i]["true_relres"] < 1e-8


@pytest.mark.parametrize("method", [VF.CG, VF.BICG, VF.BICGSTAB, VF.CORS, VF.BICORSTAB])
def test_batch_methods_shared_operator(method):
    """Several right-hand sides on one operator (vie_solve_batched) in every method."""
    sc = drude_sphere(560.0)
    rng = np.random.default_rng(7)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        B = np.stack([s.incident(0.0, 0.0, 0.0), s.incident(30.0, 20.0, 90.0),
                      rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)])
        X, infos = s.solve_batched(B, method=method, tol=1e-7)
        for k in range(B.shape[0]):
            x, info = s.solve(B[k], method=method, tol=1e-7)
            assert infos[k]["iterations"] == info["iterations"], (k, infos[k], info)
            assert rel(X[k], x) <= 1e-13, (k, rel(X[k], x))


def test_batch_fp32_context():
    sc = drude_sphere(560.0)
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP32) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        B = np.stack([s.incident(0.0, 0.0, psi) for psi in (0.0, 45.0, 90.0)])
        X, infos = s.solve_batched(B, method=VF.BICGSTAB, tol=1e-5)
        for k in range(3):
            x, info = s.solve(B[k], method=VF.BICGSTAB, tol=1e-5)
            assert infos[k]["iterations"] == info["iterations"]
            assert rel(X[k], x) <= 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_instance_at_its_limit_leaves_others_alone(mixed):
    """One instance limited to 3 iterations ends NOT_CONVERGED there; the
    others converge exactly as on their own."""
    points = [(lam, 0.0) for lam in LAMS]
    st, infos, xs, qs, ab, scenes = batch_of(points, mixed, maxits=[0, 3, 0])
    assert st == VF.NOT_CONVERGED
    assert not infos[1]["converged"] and infos[1]["iterations"] == 3
    for i in (0, 2):
        x, info, q, a = single(scenes[i], 0.0, mixed)
        assert infos[i]["converged"]
        assert infos[i]["iterations"] == info["iterations"]
        assert rel(xs[i], x) <= 1e-13


def test_batch_uneven_tolerances_compact_slots():
    """Loose and tight tolerances in one batch: the loose instances stop early
    (the launches then serve only the remaining ones) and every instance
    matches its own solve."""
    points = [(lam, psi) for lam in LAMS for psi in PSIS]
    tols = [1e-2, 1e-9, 1e-2, 1e-2, 1e-9, 1e-3]
    st, infos, xs, qs, ab, scenes = batch_of(points, False, tols=tols)
    assert st == VF.OK
    its = [inf["iterations"] for inf in infos]
    assert max(its) > 2 * min(its), its
    for i, (lam, psi) in enumerate(points):
        x, info, q, a = single(scenes[i], psi, False, tol=tols[i])
        assert infos[i]["iterations"] == info["iterations"], (i, infos[i], info)
        assert rel(xs[i], x) <= 1e-13


def test_batch_zero_rhs_is_a_status():
    """A zero right-hand side ends that instance with a breakdown status (the
    reference would loop on 0/0); the other instance solves normally."""
    sc = drude_sphere(560.0)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.batch_reserve(2)
        s.batch_set_rhs(0, np.zeros((3,) + sc.n, np.complex128))
        s.batch_set_rhs_incident(1, 0.0, 0.0, 0.0)
        st, infos = s.batch_solve(2, tol=1e-8)
        assert st == VF.BREAKDOWN
        assert not infos[0]["converged"]
        assert infos[1]["converged"]
        s.set_rhs_incident(0.0, 0.0, 0.0)
        info = s.solve_resident(VF.BICGSTAB, tol=1e-8)
        assert info["iterations"] == infos[1]["iterations"]


def test_dimer_sweep_batched_equals_sequential():
    """The 96^3 Au dimer (Drude) sweep, tolerance 1e-6, mixed precision, on a
    subset of its 64 wavelengths with both polarisations: the batched sweep
    gives every right-hand side the Q_ext, Q_abs and iteration counts of its
    own solve."""
    lams = 400.0 + 400.0 * np.arange(64) / 63.0
    points = [(float(lam), psi) for lam in lams[[0, 9, 21, 40, 63]] for psi in (0.0, 90.0)]
    scene = lambda lam, psi: G.dimer_case(lam, psi=psi)  # noqa: E731
    print("dimer sweep, batched", flush=True)
    tm = {}
    rb = SW.run_sweep_batched(scene, points, tol=1e-6, mixed=True, timings=tm)
    print("dimer sweep, sequential", tm, flush=True)
    rs = SW.run_sweep(scene, points, tol=1e-6, mixed=True)
    for a, b in zip(rb, rs):
        assert a["converged"] and b["converged"]
        assert a["iterations"] == b["iterations"], (a, b)
        assert abs(a["q_ext"] - b["q_ext"]) <= 1e-12 * abs(b["q_ext"])
        assert abs(a["q_abs"] - b["q_abs"]) <= 1e-12 * abs(b["q_abs"])
        assert a["true_relres"] < 1e-6
