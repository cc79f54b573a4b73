"""Pins the oracle's FarField2 cuts (``orc_rcs``, vie_fft/NearFar.h:302-456)
with closed forms: the far field of point current elements sitting at their
face centres (J_x of cell (i,j,k) at ((i+1)dx, (j+1/2)dy, (k+1/2)dz)), whose
amplitude along the unit direction v is dV J exp(j k0 r.v), projected on the
theta and phi unit vectors."""
import numpy as np
import pytest

from oracle import oracle as O

Z0 = 120 * np.pi


def _grid(n=(6, 5, 7), d=(0.02, 0.025, 0.03)):
    return n, d, 2 * np.pi / 5.0


def _face_centre(c, idx, d):
    p = np.array([(idx[a] + 0.5) * d[a] for a in range(3)])
    p[c] += 0.5 * d[c]
    return p


def _closed_form(elements, n, d, k0, phi_deg):
    """elements: list of (component, cell index, complex current)."""
    dV = d[0] * d[1] * d[2]
    lam = 2 * np.pi / k0
    out = []
    for plane, ph in (("theta", phi_deg), ("phi", phi_deg + 90.0)):
        vals = []
        p = np.radians(ph)
        for a in range(181):
            t = np.radians(a)
            v = np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
            S = np.zeros(3, complex)
            for c, idx, j in elements:
                S[c] += dV * j * np.exp(1j * k0 * _face_centre(c, idx, d) @ v)
            if plane == "theta":
                u = np.array([np.cos(t) * np.cos(p), np.cos(t) * np.sin(p), -np.sin(t)])
            else:
                u = np.array([-np.sin(p), np.cos(p), 0.0])
            amp = abs(u @ S) ** 2
            vals.append(10 * np.log10(k0 * k0 * Z0 * Z0 * amp / (4 * np.pi) / lam / lam) if amp > 0 else -np.inf)
        out.append(np.array(vals))
    return out


def _field(n, elements):
    J = np.zeros((3,) + tuple(n), complex)
    for c, idx, j in elements:
        J[(c,) + tuple(idx)] += j
    return J


def _check(elements, phi, n=None, d=None):
    n0, d0, k0 = _grid()
    n, d = n or n0, d or d0
    rt, rp = O.rcs(n, d, k0, _field(n, elements), 0.0, phi)
    et, ep = _closed_form(elements, n, d, k0, phi)
    for got, want in ((rt, et), (rp, ep)):
        ok = np.isfinite(want) & (want > want[np.isfinite(want)].max() - 80)  # skip pattern nulls
        assert np.max(np.abs(got[ok] - want[ok])) < 1e-8


def test_single_x_element_is_a_dipole_pattern():
    n, d, k0 = _grid()
    el = [(0, (2, 1, 3), 1.0 + 0.5j)]
    rt, rp = O.rcs(n, d, k0, _field(n, el), 0.0, 0.0)
    dV = d[0] * d[1] * d[2]
    lam = 2 * np.pi / k0
    base = k0 * k0 * Z0 * Z0 * dV * dV * abs(el[0][2]) ** 2 / (4 * np.pi) / lam / lam
    t = np.radians(np.arange(181))
    keep = np.abs(np.cos(t)) > 1e-3
    assert np.allclose(rt[keep], 10 * np.log10(base * np.cos(t[keep]) ** 2), atol=1e-9)   # E plane: cos^2
    assert np.allclose(rp, 10 * np.log10(base), atol=1e-9)                               # H plane: flat


def test_two_z_elements_array_factor():
    n, d, k0 = _grid()
    m = 4
    el = [(2, (3, 2, 1), 1.0), (2, (3, 2, 1 + m), 1.0)]
    rt, _ = O.rcs(n, d, k0, _field(n, el), 0.0, 0.0)
    dV = d[0] * d[1] * d[2]
    lam = 2 * np.pi / k0
    t = np.radians(np.arange(181))
    amp = (dV * np.sin(t)) ** 2 * 4 * np.cos(k0 * m * d[2] * np.cos(t) / 2) ** 2
    ok = amp > 1e-6 * amp.max()
    want = 10 * np.log10(k0 * k0 * Z0 * Z0 * amp[ok] / (4 * np.pi) / lam / lam)
    assert np.allclose(rt[ok], want, atol=1e-8)


@pytest.mark.parametrize("phi", [0.0, 30.0, 90.0])
def test_mixed_elements_at_face_centres(phi):
    els = [(0, (1, 2, 3), 1.0 - 0.3j), (2, (4, 1, 5), 0.4 + 0.8j), (1, (2, 3, 0), -0.7 + 0.2j)]
    _check(els, phi)
