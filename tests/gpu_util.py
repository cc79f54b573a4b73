"""Helpers shared by the GPU parity tests (GPU side: libviefft.so through the
C ABI; checker side: the CPU oracle)."""
import numpy as np

from oracle import oracle as O
from vie_fft_amd import viefft as VF


def random_medium(n, seed, active=0.85):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.5, 6.0, n) - 1j * rng.uniform(0.0, 1.0, n)
    flags = (rng.uniform(0, 1, (3,) + tuple(n)) < active).astype(np.uint8)
    return eps, flags


def random_field(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((3,) + tuple(n)) + 1j * rng.standard_normal((3,) + tuple(n))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gpu_solver(n, d, k0, eps, flags, precision=VF.FP64):
    s = VF.Solver(n, d, k0, precision)
    s.set_medium_eps(eps, flags)
    s.green_compute()
    return s


def scene_solver(sc, precision=VF.FP64):
    s = VF.Solver(sc.n, sc.d, sc.k0, precision)
    s.set_medium(sc.material(), sc.eps_table, sc.flags())
    s.green_compute()
    return s


def oracle_op(n, d, k0, eps, flags):
    return O.Operator(n, d, k0, eps, flags, O.SCHED_FUSED)
