"""Host-side interface of the gfx950 VIE-FFT library (ctypes over ``include/viefft.h``).

The object model mirrors the reference's host flow (paths relative to the
reference root): ``vie_fft/CG_FFT_VIE.cpp:52-61`` runs Parameter_Input,
GetMaterial/GetFlag, Green_Function_Compute, Incident_Wave, FFT_Program,
FarField1 and NearField on global arrays; here a :class:`Solver` owns one
grid and exposes the same steps with explicit arguments.  All compute runs
in ``libviefft.so``; loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libviefft.so")

FP64, FP32 = 0, 1
MIXED_REFINE, MIXED_RELIABLE = 0, 1
CG, BICG, BICGSTAB, CORS, BICORSTAB = 1, 2, 3, 4, 5
PRECOND_NONE, PRECOND_JACOBI = 0, 1
METHODS = {"cg": CG, "bicg": BICG, "bicgstab": BICGSTAB, "cors": CORS, "bicorstab": BICORSTAB}
OK, ERR_ARG, ERR_DEVICE, ERR_OOM, NOT_CONVERGED, BREAKDOWN, ERR_NAN, ERR_UNSUPPORTED, DIVERGED = range(9)
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_OOM", 4: "NOT_CONVERGED", 5: "BREAKDOWN",
          6: "ERR_NAN", 7: "ERR_UNSUPPORTED", 8: "DIVERGED"}

# every symbol include/viefft.h declares
EXPORTS = [
    "vie_create", "vie_destroy", "vie_last_error", "vie_supported_n", "vie_set_medium",
    "vie_set_medium_eps", "vie_green_compute", "vie_incident", "vie_apply", "vie_apply_async",
    "vie_solve", "vie_set_rhs", "vie_solve_resident", "vie_get_solution", "vie_solve_mixed_resident",
    "vie_forward_scatter", "vie_absorbed", "vie_time_kernel", "vie_kernel_time",
    "vie_model_bytes_matvec", "vie_plan", "vie_absorption", "vie_absorption_grad", "vie_absorption_grad_mixed",
    "vie_set_k0", "vie_set_rhs_incident", "vie_far_field", "vie_rcs", "vie_solve_batched",
    "vie_batch_reserve", "vie_batch_capacity", "vie_batch_set_operators", "vie_batch_set_rhs_incident",
    "vie_batch_set_rhs", "vie_batch_solve", "vie_batch_get_solution", "vie_batch_forward_scatter",
    "vie_batch_absorbed", "vie_far_field_noalloc", "vie_apply_raw", "vie_set_guess", "vie_solve_x0",
    "vie_history", "vie_set_divergence_stop", "vie_tuning", "vie_voxelize", "vie_set_medium_shapes",
    "vie_batch_outputs", "vie_set_preconditioner", "vie_release_buffers",
]


class Desc(ct.Structure):
    _fields_ = [("n", ct.c_int * 3), ("d", ct.c_double * 3), ("k0", ct.c_double), ("precision", ct.c_int)]


class Info(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("converged", ct.c_int), ("err", ct.c_double),
                ("true_relres", ct.c_double), ("outer", ct.c_int), ("seconds", ct.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SPHERE, ZBAND = 0, 1


class Shape(ct.Structure):
    """vie_shape: a sphere (centre, radius) or a z band [z0, z1), in cells."""
    _fields_ = [("kind", ct.c_int), ("mat", ct.c_int), ("c", ct.c_double * 3), ("r", ct.c_double),
                ("z0", ct.c_double), ("z1", ct.c_double)]


def shapes_of(scene):
    """The vie_shape array of a geometry.Scene's shapes (spheres and z bands)."""
    from . import geometry as G

    arr = (Shape * max(len(scene.shapes), 1))()
    for i, sh in enumerate(scene.shapes):
        if isinstance(sh, G.Sphere):
            arr[i].kind, arr[i].mat, arr[i].r = SPHERE, int(sh.mat), float(sh.radius)
            for a in range(3):
                arr[i].c[a] = float(sh.center[a])
        elif isinstance(sh, G.Band):
            arr[i].kind, arr[i].mat, arr[i].z0, arr[i].z1 = ZBAND, int(sh.mat), float(sh.z0), float(sh.z1)
        else:
            raise TypeError(sh)
    return arr, len(scene.shapes)


class VieError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load():
    """Load libviefft.so (built by ``vie_fft_amd.build``); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run vie_fft_amd/build.py (hipcc, gfx950) first")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7);
    # only the first one loaded is used by the process, and torch cannot run
    # on another release's runtime.  Load torch's first when it is installed
    # so that torch and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(LIB_PATH)
    dp = ct.POINTER(ct.c_double)
    vp = ct.c_void_p
    ub = ct.POINTER(ct.c_ubyte)
    L.vie_create.argtypes = [ct.POINTER(vp), ct.POINTER(Desc)]
    L.vie_destroy.argtypes = [vp]
    L.vie_last_error.argtypes = [vp]
    L.vie_last_error.restype = ct.c_char_p
    L.vie_supported_n.argtypes = [ct.c_int]
    L.vie_set_medium.argtypes = [vp, ub, dp, ct.c_int, ub]
    L.vie_set_medium_eps.argtypes = [vp, dp, ub]
    L.vie_green_compute.argtypes = [vp]
    L.vie_incident.argtypes = [vp, ct.c_double, ct.c_double, ct.c_double, dp]
    L.vie_apply.argtypes = [vp, dp, dp, ct.c_int]
    L.vie_apply_raw.argtypes = [vp, dp, dp, ct.c_int]
    L.vie_apply_async.argtypes = [vp, vp, vp, ct.c_int, vp]
    L.vie_solve.argtypes = [vp, ct.c_int, ct.c_double, ct.c_int, dp, dp, ct.POINTER(Info)]
    L.vie_set_rhs.argtypes = [vp, dp]
    L.vie_solve_resident.argtypes = [vp, ct.c_int, ct.c_double, ct.c_int, ct.POINTER(Info)]
    L.vie_get_solution.argtypes = [vp, dp]
    L.vie_set_guess.argtypes = [vp, dp]
    L.vie_history.argtypes = [vp, ct.c_int, dp, ct.POINTER(ct.c_int)]
    L.vie_set_divergence_stop.argtypes = [vp, ct.c_double]
    L.vie_set_preconditioner.argtypes = [vp, ct.c_int]
    L.vie_release_buffers.argtypes = [ct.POINTER(ct.c_size_t)]
    L.vie_tuning.argtypes = [vp, ct.POINTER(ct.c_int), dp]
    L.vie_voxelize.argtypes = [vp, ct.c_int, ct.POINTER(Shape), ub, ub]
    L.vie_set_medium_shapes.argtypes = [vp, ct.c_int, ct.POINTER(Shape), dp, ct.c_int]
    L.vie_solve_x0.argtypes = [vp, ct.c_int, ct.c_double, ct.c_int, dp, dp, ct.POINTER(Info)]
    L.vie_solve_mixed_resident.argtypes = [vp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_int,
                                           ct.POINTER(Info)]
    L.vie_forward_scatter.argtypes = [vp, ct.c_double, ct.c_double, ct.c_double, dp]
    L.vie_absorbed.argtypes = [vp, ct.c_int, dp]
    L.vie_time_kernel.argtypes = [vp, ct.c_int]
    L.vie_kernel_time.argtypes = [vp, ct.POINTER(ct.c_int), dp]
    L.vie_model_bytes_matvec.argtypes = [vp]
    L.vie_model_bytes_matvec.restype = ct.c_double
    L.vie_plan.argtypes = [vp, ct.POINTER(ct.c_int)]
    L.vie_absorption.argtypes = [vp, dp]
    L.vie_set_k0.argtypes = [vp, ct.c_double]
    L.vie_set_rhs_incident.argtypes = [vp, ct.c_double, ct.c_double, ct.c_double]
    L.vie_far_field.argtypes = [vp, ct.c_int, dp, dp, dp]
    L.vie_far_field_noalloc.argtypes = [vp, ct.c_int, dp, dp, dp]
    L.vie_rcs.argtypes = [vp, ct.c_double, dp, dp]
    L.vie_solve_batched.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_int, dp, dp,
                                    ct.POINTER(Info)]
    L.vie_absorption_grad.argtypes = [vp, ct.c_int, ct.c_double, ct.c_int, dp, dp, dp, ct.POINTER(Info)]
    L.vie_absorption_grad_mixed.argtypes = [vp, ct.c_double, ct.c_int, dp, dp, dp, ct.POINTER(Info)]
    ip = ct.POINTER(ct.c_int)
    L.vie_batch_reserve.argtypes = [vp, ct.c_int]
    L.vie_batch_capacity.argtypes = [vp, ip]
    L.vie_batch_set_operators.argtypes = [vp, ct.c_int, ct.c_int, dp, dp, ct.c_int]
    L.vie_batch_set_rhs_incident.argtypes = [vp, ct.c_int, ct.c_double, ct.c_double, ct.c_double]
    L.vie_batch_set_rhs.argtypes = [vp, ct.c_int, dp]
    L.vie_batch_solve.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_int, dp, ip, ct.POINTER(Info)]
    L.vie_batch_get_solution.argtypes = [vp, ct.c_int, dp]
    L.vie_batch_forward_scatter.argtypes = [vp, ct.c_int, ct.c_double, ct.c_double, ct.c_double, dp]
    L.vie_batch_absorbed.argtypes = [vp, ct.c_int, ct.c_int, dp]
    L.vie_batch_outputs.argtypes = [vp, ct.c_int, dp, dp, dp, ct.c_int, dp, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _ub(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_ubyte))


def release_buffers():
    """Hand the device buffers the library keeps for reuse back to the HIP
    runtime (vie_release_buffers); returns the bytes released."""
    n = ct.c_size_t()
    st = load().vie_release_buffers(ct.byref(n))
    if st != OK:
        raise VieError(st, "vie_release_buffers failed")
    return n.value


class Solver:
    """One VIE grid on the GPU: operator, Green's function, solves, post-processing."""

    def __init__(self, n, d, k0, precision=FP64):
        self.lib = load()
        self.n = tuple(int(v) for v in n)
        self.V = int(np.prod(self.n))
        desc = Desc()
        for a in range(3):
            desc.n[a] = self.n[a]
            desc.d[a] = float(d[a])
        desc.k0 = float(k0)
        desc.precision = int(precision)
        self.d = tuple(float(v) for v in d)
        self.k0 = float(k0)
        self.precision = int(precision)
        h = ct.c_void_p()
        st = self.lib.vie_create(ct.byref(h), ct.byref(desc))
        if st != OK:
            raise VieError(st, "vie_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.vie_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st, allow=(OK,)):
        if st not in allow:
            msg = self.lib.vie_last_error(self.h)
            raise VieError(st, msg.decode() if msg else "")
        return st

    def _field(self, x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape((3,) + self.n))

    # -- medium and set-up -------------------------------------------------
    def set_medium(self, material, eps_table, flags):
        """Material ids (V bytes), eps per id (engineering sign), flags (3,V)."""
        m = np.ascontiguousarray(material, dtype=np.uint8).reshape(self.n)
        e = np.ascontiguousarray(eps_table, dtype=np.complex128).ravel()
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape((3,) + self.n)
        self._check(self.lib.vie_set_medium(self.h, _ub(m), _dp(e), e.size, _ub(f)))

    def set_medium_eps(self, eps, flags):
        e = np.ascontiguousarray(eps, dtype=np.complex128).reshape(self.n)
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape((3,) + self.n)
        self._check(self.lib.vie_set_medium_eps(self.h, _dp(e), _ub(f)))

    def voxelize(self, scene):
        """(material, flags) of the scene's shapes, computed on the device."""
        arr, n = shapes_of(scene)
        m = np.zeros(self.n, np.uint8)
        f = np.zeros((3,) + self.n, np.uint8)
        self._check(self.lib.vie_voxelize(self.h, n, arr, _ub(m), _ub(f)))
        return m, f

    def set_medium_shapes(self, scene, eps_table=None):
        """Medium from the scene's shapes (voxelised on the device) and its eps table."""
        arr, n = shapes_of(scene)
        e = np.ascontiguousarray(scene.eps_table if eps_table is None else eps_table, dtype=np.complex128).ravel()
        self._check(self.lib.vie_set_medium_shapes(self.h, n, arr, _dp(e), e.size))

    def set_k0(self, k0):
        """Move the context to another wavenumber (then green_compute, set_rhs)."""
        self._check(self.lib.vie_set_k0(self.h, float(k0)))
        self.k0 = float(k0)

    def green_compute(self):
        """Green_Function_Compute (vie_fft/GreenFunction.h:8-141)."""
        self._check(self.lib.vie_green_compute(self.h))

    def incident(self, theta=0.0, phi=0.0, psi=0.0):
        """Incident_Wave (vie_fft/GreenFunction.h:144-248), unmasked."""
        E = np.zeros((3,) + self.n, np.complex128)
        self._check(self.lib.vie_incident(self.h, theta, phi, psi, _dp(E)))
        return E

    # -- operator ----------------------------------------------------------
    def apply(self, x, adjoint=False):
        """y = A x (X_FFT/Y_FFT/Z_FFT + Current_Reset) or A^H x (the _Conjugate set)."""
        x = self._field(x)
        y = np.zeros_like(x)
        self._check(self.lib.vie_apply(self.h, _dp(x), _dp(y), int(adjoint)))
        return y

    def apply_raw(self, x, adjoint=False):
        """A x or A^H x without the output mask (X_FFT/Y_FFT/Z_FFT before Current_Reset)."""
        x = self._field(x)
        y = np.zeros_like(x)
        self._check(self.lib.vie_apply_raw(self.h, _dp(x), _dp(y), int(adjoint)))
        return y

    def apply_async(self, x_ptr, y_ptr, adjoint=False, stream=None):
        """Device-pointer product (e.g. torch tensors' data_ptr()), enqueued on stream."""
        self._check(self.lib.vie_apply_async(self.h, ct.c_void_p(x_ptr), ct.c_void_p(y_ptr), int(adjoint),
                                             ct.c_void_p(stream) if stream else None))

    # -- solves --------------------------------------------------------------
    def solve(self, b, method=BICGSTAB, tol=1e-3, maxit=0):
        """FFT_Program (vie_fft/CG_Iteration.h:1365-1388): returns (x, info)."""
        method = METHODS.get(method, method) if isinstance(method, str) else method
        b = self._field(b)
        x = np.zeros_like(b)
        info = Info()
        self._check(self.lib.vie_solve(self.h, int(method), float(tol), int(maxit), _dp(b), _dp(x),
                                       ct.byref(info)), allow=(OK, NOT_CONVERGED))
        return x, info.as_dict()

    def solve_x0(self, b, x0, method=BICGSTAB, tol=1e-3, maxit=0):
        """vie_solve from the initial guess x0: returns (x, info)."""
        method = METHODS.get(method, method) if isinstance(method, str) else method
        b = self._field(b)
        x = np.array(self._field(x0), copy=True)
        info = Info()
        self._check(self.lib.vie_solve_x0(self.h, int(method), float(tol), int(maxit), _dp(b), _dp(x),
                                          ct.byref(info)), allow=(OK, NOT_CONVERGED))
        return x, info.as_dict()

    def set_guess(self, x0=None):
        """Initial guess of the next resident solve: x0, or the resident solution (None)."""
        if x0 is None:
            self._check(self.lib.vie_set_guess(self.h, None))
        else:
            x = self._field(x0)
            self._check(self.lib.vie_set_guess(self.h, _dp(x)))

    def solve_batched(self, B, method=BICGSTAB, tol=1e-3, maxit=0, mixed=False):
        """Solve for several right-hand sides B (nrhs, 3, Nx, Ny, Nz): returns (X, infos)."""
        method = METHODS.get(method, method) if isinstance(method, str) else method
        B = np.ascontiguousarray(np.asarray(B, dtype=np.complex128).reshape((-1, 3) + self.n))
        X = np.zeros_like(B)
        infos = (Info * max(B.shape[0], 1))()
        self._check(self.lib.vie_solve_batched(self.h, B.shape[0], int(bool(mixed)), int(method), float(tol),
                                               int(maxit), _dp(B), _dp(X), infos),
                    allow=(OK, NOT_CONVERGED, BREAKDOWN))
        return X, [infos[k].as_dict() for k in range(B.shape[0])]

    def set_rhs(self, b):
        b = self._field(b)
        self._check(self.lib.vie_set_rhs(self.h, _dp(b)))

    def set_rhs_incident(self, theta=0.0, phi=0.0, psi=0.0):
        """Resident right-hand side = masked incident field, computed on the device."""
        self._check(self.lib.vie_set_rhs_incident(self.h, float(theta), float(phi), float(psi)))

    def solve_resident(self, method=BICGSTAB, tol=1e-3, maxit=0):
        method = METHODS.get(method, method) if isinstance(method, str) else method
        info = Info()
        self._check(self.lib.vie_solve_resident(self.h, int(method), float(tol), int(maxit), ct.byref(info)),
                    allow=(OK, NOT_CONVERGED))
        return info.as_dict()

    def solve_mixed(self, tol=1e-6, inner_param=0.0, maxit=0, inner_method=BICGSTAB, strategy=MIXED_RELIABLE):
        """FP32 Krylov recurrences with FP64 solution and FP64 true-residual stop rule
        (resident b -> x).  strategy MIXED_RELIABLE: residual replacement every
        inner_param reduction; MIXED_REFINE: restarted refinement, inner tolerance inner_param."""
        inner_method = METHODS.get(inner_method, inner_method) if isinstance(inner_method, str) else inner_method
        info = Info()
        self._check(self.lib.vie_solve_mixed_resident(self.h, int(strategy), int(inner_method), float(tol),
                                                      float(inner_param), int(maxit), ct.byref(info)),
                    allow=(OK, NOT_CONVERGED))
        return info.as_dict()

    def solution(self):
        x = np.zeros((3,) + self.n, np.complex128)
        self._check(self.lib.vie_get_solution(self.h, _dp(x)))
        return x

    # -- post-processing -----------------------------------------------------
    def forward_scatter(self, theta=0.0, phi=0.0, psi=0.0):
        """Forward_RCS of FarField1 (vie_fft/NearFar.h:16-109) for the resident solution."""
        v = ct.c_double()
        self._check(self.lib.vie_forward_scatter(self.h, theta, phi, psi, ct.byref(v)))
        return v.value

    def far_field(self, theta_deg, phi_deg):
        """(ndir, 3) complex far-field amplitudes of the resident solution."""
        th = np.ascontiguousarray(theta_deg, dtype=np.float64).ravel()
        ph = np.ascontiguousarray(phi_deg, dtype=np.float64).ravel()
        if th.shape != ph.shape:
            raise ValueError("theta and phi must have the same length")
        S = np.zeros((th.size, 3), np.complex128)
        self._check(self.lib.vie_far_field(self.h, th.size, _dp(th), _dp(ph), _dp(S)))
        return S

    def far_field_noalloc(self, theta_deg, phi_deg):
        """far_field through the entry point that allocates no device memory
        (at most 512 directions per call)."""
        th = np.ascontiguousarray(theta_deg, dtype=np.float64).ravel()
        ph = np.ascontiguousarray(phi_deg, dtype=np.float64).ravel()
        if th.shape != ph.shape:
            raise ValueError("theta and phi must have the same length")
        S = np.zeros((th.size, 3), np.complex128)
        self._check(self.lib.vie_far_field_noalloc(self.h, th.size, _dp(th), _dp(ph), _dp(S)))
        return S

    def rcs(self, phi=0.0):
        """FarField2 (vie_fft/NearFar.h:302-456): (E-plane, H-plane) RCS in dB, 181 angles each."""
        rt = np.zeros(181)
        rp = np.zeros(181)
        self._check(self.lib.vie_rcs(self.h, float(phi), _dp(rt), _dp(rp)))
        return rt, rp

    def absorbed(self, mat):
        """Ene_Metal of NearField (vie_fft/NearFar.h:180-250) for material id `mat`."""
        v = ct.c_double()
        self._check(self.lib.vie_absorbed(self.h, int(mat), ct.byref(v)))
        return v.value

    def absorption(self):
        """Absorbed power of the resident solution over all active faces."""
        v = ct.c_double()
        self._check(self.lib.vie_absorption(self.h, ct.byref(v)))
        return v.value

    def absorption_grad(self, method=BICGSTAB, tol=1e-10, maxit=0, mixed=False):
        """(P, dP/dRe eps, dP/dIm eps, info of the adjoint solve) for the resident solution;
        mixed: the adjoint solve by the mixed-precision BiCGSTAB (FP64 contexts)."""
        method = METHODS.get(method, method) if isinstance(method, str) else method
        v = ct.c_double()
        gre = np.zeros(self.n)
        gim = np.zeros(self.n)
        info = Info()
        if mixed:
            st = self.lib.vie_absorption_grad_mixed(self.h, float(tol), int(maxit), ct.byref(v), _dp(gre), _dp(gim),
                                                    ct.byref(info))
        else:
            st = self.lib.vie_absorption_grad(self.h, int(method), float(tol), int(maxit), ct.byref(v), _dp(gre),
                                              _dp(gim), ct.byref(info))
        self._check(st, allow=(OK, NOT_CONVERGED))
        return v.value, gre, gim, info.as_dict()

    # -- batches of instances (one geometry, several wavelengths / polarisations)
    def batch_reserve(self, nb):
        """Device storage for nb instances solved together (one operator for all)."""
        self._check(self.lib.vie_batch_reserve(self.h, int(nb)))

    def batch_capacity(self):
        v = ct.c_int()
        self._check(self.lib.vie_batch_capacity(self.h, ct.byref(v)))
        return v.value

    def batch_set_operators(self, i0, k0, eps_tables=None):
        """Own operators for instances i0.. : wavenumbers k0 (count,) and eps
        tables (count, n_mat) complex, or None for the context's permittivity."""
        k = np.ascontiguousarray(np.atleast_1d(k0), dtype=np.float64)
        if eps_tables is None:
            self._check(self.lib.vie_batch_set_operators(self.h, int(i0), k.size, _dp(k), None, 0))
        else:
            e = np.ascontiguousarray(np.asarray(eps_tables, dtype=np.complex128).reshape(k.size, -1))
            self._check(self.lib.vie_batch_set_operators(self.h, int(i0), k.size, _dp(k), _dp(e), e.shape[1]))

    def batch_set_rhs_incident(self, i, theta=0.0, phi=0.0, psi=0.0):
        self._check(self.lib.vie_batch_set_rhs_incident(self.h, int(i), float(theta), float(phi), float(psi)))

    def batch_set_rhs(self, i, b):
        b = self._field(b)
        self._check(self.lib.vie_batch_set_rhs(self.h, int(i), _dp(b)))

    def batch_solve(self, nb, method=BICGSTAB, tol=1e-6, maxit=0, mixed=False, tols=None, maxits=None):
        """Solve instances 0..nb-1 together; returns (status, infos)."""
        method = METHODS.get(method, method) if isinstance(method, str) else method
        infos = (Info * nb)()
        tv = None if tols is None else np.ascontiguousarray(tols, dtype=np.float64)
        mv = None if maxits is None else np.ascontiguousarray(maxits, dtype=np.int32)
        st = self.lib.vie_batch_solve(self.h, int(nb), int(bool(mixed)), int(method), float(tol), int(maxit),
                                      _dp(tv) if tv is not None else None,
                                      mv.ctypes.data_as(ct.POINTER(ct.c_int)) if mv is not None else None, infos)
        self._check(st, allow=(OK, NOT_CONVERGED, BREAKDOWN))
        return st, [infos[k].as_dict() for k in range(nb)]

    def batch_solution(self, i):
        x = np.zeros((3,) + self.n, np.complex128)
        self._check(self.lib.vie_batch_get_solution(self.h, int(i), _dp(x)))
        return x

    def batch_forward_scatter(self, i, theta=0.0, phi=0.0, psi=0.0):
        v = ct.c_double()
        self._check(self.lib.vie_batch_forward_scatter(self.h, int(i), theta, phi, psi, ct.byref(v)))
        return v.value

    def batch_absorbed(self, i, mat):
        v = ct.c_double()
        self._check(self.lib.vie_batch_absorbed(self.h, int(i), int(mat), ct.byref(v)))
        return v.value

    def batch_outputs(self, theta, phi, psi, mat=None):
        """Forward scatter (and, with mat, absorbed power) of instances 0..len(theta)-1
        with one read-back: (forward, absorbed or None) arrays."""
        th, ph, ps = (np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()) for a in (theta, phi, psi))
        nb = th.size
        if ph.size != nb or ps.size != nb:
            raise ValueError("theta, phi and psi need one angle per instance")
        fw = np.zeros(nb)
        ab = np.zeros(nb) if mat is not None else None
        self._check(self.lib.vie_batch_outputs(self.h, nb, _dp(th), _dp(ph), _dp(ps), int(mat or 0), _dp(fw),
                                               _dp(ab) if ab is not None else None))
        return fw, ab

    # -- instrumentation -----------------------------------------------------
    def tuning(self):
        """The autotuner's plan per precision: tuned, variant bits, x/y/z pencil lengths, us per product."""
        a = (ct.c_int * 10)()
        u = (ct.c_double * 4)()
        self._check(self.lib.vie_tuning(self.h, a, u))
        out = {}
        for k, name in enumerate(("fp64", "fp32")):
            o = a[5 * k: 5 * k + 5]
            if o[0] >= 0:
                out[name] = {"tuned": bool(o[0]), "variant": o[1], "pencils": tuple(o[2:5]),
                             "us_per_product": u[k], "us_per_product_defaults": u[k + 2]}
        return out

    def set_divergence_stop(self, factor):
        """Later solves end with DIVERGED once the stop quantity exceeds factor x its minimum (0: off)."""
        self._check(self.lib.vie_set_divergence_stop(self.h, float(factor)))

    def set_preconditioner(self, kind):
        """Later BiCGSTAB solves use the Jacobi preconditioner (PRECOND_JACOBI) or none
        (PRECOND_NONE, the reference's iterates)."""
        self._check(self.lib.vie_set_preconditioner(self.h, int(kind)))

    def history(self, max_n=65536):
        """Stop quantity after each iteration of the last single solve."""
        buf = np.zeros(max_n)
        cnt = ct.c_int()
        self._check(self.lib.vie_history(self.h, int(max_n), _dp(buf), ct.byref(cnt)))
        return buf[: min(cnt.value, max_n)].copy()

    def time_kernel(self, cls):
        self._check(self.lib.vie_time_kernel(self.h, int(cls)))

    def kernel_time(self):
        n = ct.c_int()
        ms = ct.c_double()
        self._check(self.lib.vie_kernel_time(self.h, ct.byref(n), ct.byref(ms)))
        return n.value, ms.value

    def model_bytes_matvec(self):
        return self.lib.vie_model_bytes_matvec(self.h)

    def plan(self):
        """Internal plan of the resident medium: axis order, x window of the
        products inside solves, pencil length of their x pass, internal grid,
        y and z windows and pencil lengths of the products inside solves."""
        a = (ct.c_int * 15)()
        self._check(self.lib.vie_plan(self.h, a))
        return {"axis_order": tuple(a[0:3]), "x_window": (a[3], a[4]), "x_pencil": a[5], "grid": tuple(a[6:9]),
                "y_window": (a[9], a[10]), "y_pencil": a[11], "z_window": (a[12], a[13]), "z_pencil": a[14]}


def supported_n(n):
    return bool(load().vie_supported_n(int(n)))
