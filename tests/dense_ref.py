"""Independent dense real-space assembly of the VIE operator for tiny grids.

Test helper (no FFT): A[f, n; s, n'] = K_fs(n - n') with
  K_ff(m) = j k0 Z0 d_f g(m) + j Z0/(k0 d_f) (g(m+e_f) + g(m-e_f) - 2 g(m))
  K_fs(m) = j Z0/(k0 d_s) (g(m+e_f) - g(m) - g(m+e_f-e_s) + g(m-e_s)),  f != s
acting on sources with their own-axis last plane removed, plus the diagonal
face term and the output mask.  In Fourier space these are the symbols of
vie_fft/fft.h:76-100 and :256-281 (a factor exp(+j theta_a) is a shift by +1
along axis a); g is the closed form of vie_fft/GreenFunction.h:59-85, pinned
against quadrature in test_oracle_green.py.
"""
import numpy as np

from oracle import oracle as O

Z0 = 120 * np.pi


def dense_operator(n, d, k0, eps, flags):
    n = tuple(n)
    V = int(np.prod(n))
    goff = O.green_offsets(n, d, k0)

    def g(m):
        return goff[np.abs(m[..., 0]), np.abs(m[..., 1]), np.abs(m[..., 2])]

    idx = np.stack(np.meshgrid(*[np.arange(v) for v in n], indexing="ij"), -1).reshape(-1, 3)
    M = idx[:, None, :] - idx[None, :, :]
    A = np.zeros((3 * V, 3 * V), complex)
    E = np.eye(3, dtype=int)
    for f in range(3):
        for s in range(3):
            if f == s:
                K = 1j * k0 * Z0 * d[f] * g(M) + 1j * Z0 / (k0 * d[f]) * (g(M + E[f]) + g(M - E[f]) - 2 * g(M))
            else:
                K = 1j * Z0 / (k0 * d[s]) * (g(M + E[f]) - g(M) - g(M + E[f] - E[s]) + g(M - E[s]))
            keep = idx[:, s] < n[s] - 1
            A[f * V:(f + 1) * V, s * V:(s + 1) * V] = K * keep[None, :]
    dd = diag_terms(n, d, k0, eps).reshape(3, V)
    for f in range(3):
        A[f * V:(f + 1) * V, f * V:(f + 1) * V] += np.diag(dd[f])
    fl = np.asarray(flags).reshape(3 * V).astype(float)
    return fl[:, None] * A


def diag_terms(n, d, k0, eps):
    """-j (d_f/2)(Z0/k0)(1/(eps[i+e_f]-1) + 1/(eps[i]-1)) for i_f < N_f - 1."""
    eps = np.asarray(eps).reshape(n)
    out = np.zeros((3,) + tuple(n), complex)
    chi = 1.0 / (eps - 1.0)
    for f in range(3):
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[f] = slice(0, n[f] - 1)
        sl_hi[f] = slice(1, n[f])
        out[(f,) + tuple(sl_lo)] = -1j * (d[f] / 2) * Z0 / k0 * (chi[tuple(sl_hi)] + chi[tuple(sl_lo)])
    return out


def own_axis_flags(n):
    """All faces active except each component's last plane along its own axis."""
    fl = np.ones((3,) + tuple(n), np.uint8)
    for f in range(3):
        sl = [f, slice(None), slice(None), slice(None)]
        sl[1 + f] = n[f] - 1
        fl[tuple(sl)] = 0
    return fl
