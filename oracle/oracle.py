"""ctypes front end of the CPU FP64 oracle (``libvie_oracle.so``).

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the bench's cpu_baseline
leg use this module as the checker.  The product library never imports it.

Arrays are numpy complex128 in the reference layout: C-order ``[i][j][k]``,
fields are ``(3, Nx, Ny, Nz)`` (components x|y|z back to back).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VIE_ORACLE_LIB selects another build of the same source (the sanitizer build)
_LIB_PATH = os.environ.get("VIE_ORACLE_LIB") or os.path.join(_HERE, "libvie_oracle.so")
_lib = None


class Grid(ct.Structure):
    _fields_ = [("n", ct.c_int * 3), ("d", ct.c_double * 3), ("k0", ct.c_double)]


class Info(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("converged", ct.c_int), ("err", ct.c_double)]


MATVEC = ct.CFUNCTYPE(None, ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.c_int)

SCHED_REFERENCE = 0
SCHED_FUSED = 1


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        dp = ct.POINTER(ct.c_double)
        L.orc_fft1d.argtypes = [dp, ct.c_int, ct.c_int]
        L.orc_fft3d.argtypes = [dp, ct.c_int, ct.c_int, ct.c_int, ct.c_int]
        L.orc_green_offsets.argtypes = [ct.POINTER(Grid), dp]
        L.orc_green.argtypes = [ct.POINTER(Grid), dp]
        L.orc_incident.argtypes = [ct.POINTER(Grid), ct.c_double, ct.c_double, ct.c_double, dp]
        L.orc_diag.argtypes = [ct.POINTER(Grid), dp, dp, ct.c_int]
        L.orc_op_create.argtypes = [ct.POINTER(Grid), dp, dp, ct.POINTER(ct.c_ubyte), ct.c_int]
        L.orc_op_create.restype = ct.c_void_p
        L.orc_op_apply.argtypes = [ct.c_void_p, dp, dp, ct.c_int]
        L.orc_op_destroy.argtypes = [ct.c_void_p]
        L.orc_op_time_component.argtypes = [ct.c_void_p, dp, dp, ct.c_int]
        L.orc_op_time_component.restype = ct.c_double
        L.orc_solve.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_long, dp, dp,
                                ct.c_double, ct.c_int, dp, ct.POINTER(Info)]
        L.orc_forward_scatter.argtypes = [ct.POINTER(Grid), dp, ct.c_double, ct.c_double, ct.c_double]
        L.orc_forward_scatter.restype = ct.c_double
        L.orc_absorbed.argtypes = [ct.POINTER(Grid), dp, dp, ct.POINTER(ct.c_ubyte), ct.c_int]
        L.orc_absorbed.restype = ct.c_double
        L.orc_rcs.argtypes = [ct.POINTER(Grid), dp, ct.c_double, ct.c_double, dp, dp]
        ub = ct.POINTER(ct.c_ubyte)
        L.orc_absorption.argtypes = [ct.POINTER(Grid), dp, dp, ub]
        L.orc_absorption.restype = ct.c_double
        L.orc_absorption_grad.argtypes = [ct.POINTER(Grid), ct.c_void_p, dp, dp, ub, ct.c_int, ct.c_double,
                                          ct.c_int, dp, dp, ct.POINTER(Info)]
        L.orc_absorption_grad.restype = ct.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def grid(n, d, k0):
    g = Grid()
    for a in range(3):
        g.n[a] = int(n[a])
        g.d[a] = float(d[a])
    g.k0 = float(k0)
    return g


def fft1d(a, sign=-1):
    a = _c128(a).copy()
    lib().orc_fft1d(_dp(a), a.size, sign)
    return a


def fft3d(a, sign=-1):
    a = _c128(a).copy()
    lib().orc_fft3d(_dp(a), a.shape[0], a.shape[1], a.shape[2], sign)
    return a


def green_offsets(n, d, k0):
    out = np.zeros((n[0] + 1, n[1] + 1, n[2] + 1), np.complex128)
    lib().orc_green_offsets(ct.byref(grid(n, d, k0)), _dp(out))
    return out


def green(n, d, k0):
    out = np.zeros((2 * n[0], 2 * n[1], 2 * n[2]), np.complex128)
    lib().orc_green(ct.byref(grid(n, d, k0)), _dp(out))
    return out


def incident(n, d, k0, theta=0.0, phi=0.0, psi=0.0):
    out = np.zeros((3, n[0], n[1], n[2]), np.complex128)
    lib().orc_incident(ct.byref(grid(n, d, k0)), theta, phi, psi, _dp(out))
    return out


def diag(n, d, k0, eps, adjoint=False):
    eps = _c128(eps)
    out = np.zeros((3, n[0], n[1], n[2]), np.complex128)
    lib().orc_diag(ct.byref(grid(n, d, k0)), _dp(eps), _dp(out), int(adjoint))
    return out


class Operator:
    """The VIE operator y = A x on one grid (``orc_op``)."""

    def __init__(self, n, d, k0, eps, flags, schedule=SCHED_FUSED, G=None):
        self.n = tuple(int(v) for v in n)
        self.g = grid(n, d, k0)
        self.G = _c128(green(n, d, k0) if G is None else G)
        self.eps = _c128(eps).reshape(self.n)
        self.flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape((3,) + self.n)
        self.h = lib().orc_op_create(ct.byref(self.g), _dp(self.G), _dp(self.eps),
                                     self.flags.ctypes.data_as(ct.POINTER(ct.c_ubyte)), schedule)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_op_destroy(self.h)
            self.h = None

    def apply(self, x, adjoint=False):
        x = _c128(x).reshape((3,) + self.n)
        y = np.zeros_like(x)
        lib().orc_op_apply(self.h, _dp(x), _dp(y), int(adjoint))
        return y

    def time_component(self, x, f=0):
        """Seconds for one output component of the product (reference schedule work)."""
        x = _c128(x).reshape((3,) + self.n)
        y = np.zeros_like(x)
        return lib().orc_op_time_component(self.h, _dp(x), _dp(y), int(f))

    def solve(self, b, method=3, tol=1e-6, maxit=None, history=False):
        b = _c128(b).reshape((3,) + self.n)
        return solve(method, None, b, tol=tol, maxit=maxit, history=history, op=self)


def solve(method, matvec, b, tol=1e-6, maxit=None, history=False, op=None):
    """Run one of the five Krylov loops.  ``matvec(x, adjoint) -> y`` works on
    flat complex vectors; with ``op`` the oracle operator is called directly."""
    b = _c128(b)
    n = b.size
    if maxit is None:
        maxit = n if op is None else int(np.prod(op.n))
    x = np.zeros_like(b)
    hist = np.zeros(max(maxit, 1), np.float64)
    info = Info()
    if op is not None:
        fn = ct.cast(lib().orc_op_matvec, ct.c_void_p)
        ctx = op.h
        keep = None
    else:
        def _cb(_ctx, xp, yp, adj):
            xv = np.ctypeslib.as_array(xp, shape=(2 * n,)).view(np.complex128)
            yv = np.ctypeslib.as_array(yp, shape=(2 * n,)).view(np.complex128)
            yv[:] = matvec(xv.copy(), bool(adj))
        keep = MATVEC(_cb)
        fn = ct.cast(keep, ct.c_void_p)
        ctx = None
    lib().orc_solve(int(method), fn, ctx, n, _dp(b), _dp(x), float(tol), int(maxit), _dp(hist),
                    ct.byref(info))
    del keep
    out = {"x": x, "iterations": info.iterations, "converged": bool(info.converged), "err": info.err}
    if history:
        out["history"] = hist[: info.iterations].copy()
    return out


def forward_scatter(n, d, k0, J, theta=0.0, phi=0.0, psi=0.0):
    J = _c128(J)
    return lib().orc_forward_scatter(ct.byref(grid(n, d, k0)), _dp(J), theta, phi, psi)


def absorbed(n, d, k0, J, eps, material, mat_id):
    J = _c128(J)
    eps = _c128(eps)
    m = np.ascontiguousarray(material, dtype=np.uint8)
    return lib().orc_absorbed(ct.byref(grid(n, d, k0)), _dp(J), _dp(eps),
                              m.ctypes.data_as(ct.POINTER(ct.c_ubyte)), int(mat_id))


def rcs(n, d, k0, J, theta=0.0, phi=0.0):
    J = _c128(J)
    rt = np.zeros(181)
    rp = np.zeros(181)
    lib().orc_rcs(ct.byref(grid(n, d, k0)), _dp(J), theta, phi, _dp(rt), _dp(rp))
    return rt, rp


def absorption(n, d, k0, J, eps, flags):
    """Absorbed power over all active faces (``orc_absorption``)."""
    J = _c128(J)
    eps = _c128(eps)
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    return lib().orc_absorption(ct.byref(grid(n, d, k0)), _dp(J), _dp(eps),
                                f.ctypes.data_as(ct.POINTER(ct.c_ubyte)))


def absorption_grad(op, J, method=3, tol=1e-12, maxit=None):
    """(P, dP/dRe eps, dP/dIm eps, info) for the solution J of op (``orc_absorption_grad``)."""
    J = _c128(J).reshape((3,) + op.n)
    gre = np.zeros(op.n)
    gim = np.zeros(op.n)
    info = Info()
    P = lib().orc_absorption_grad(ct.byref(op.g), op.h, _dp(J), _dp(op.eps),
                                  op.flags.ctypes.data_as(ct.POINTER(ct.c_ubyte)), int(method), float(tol),
                                  int(maxit if maxit else 3 * int(np.prod(op.n))), _dp(gre), _dp(gim),
                                  ct.byref(info))
    return P, gre, gim, {"iterations": info.iterations, "converged": bool(info.converged), "err": info.err}
