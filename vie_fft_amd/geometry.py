"""Voxel media and incidence for the benchmark and test workloads.

Rules follow the reference (paths relative to its root):
  * materials are sampled at cell centres ((i+1/2)d, (j+1/2)d, (k+1/2)d)
    (GetMaterial, vie_fft/Model.h:305-332);
  * Flag_f is set when the face centre and the two points half a cell on either
    side of it along f are all non-background (GetFlag, vie_fft/Model.h:335-398);
    points outside the box count as background, so a component's last plane
    along its own axis is never active;
  * background is material 0 with the dummy permittivity 1.1 (vie_fft/Model.h:82);
  * a sphere holds a point when dist <= R (vie_fft/Model.h:229); overlapping
    shapes take the first one that holds the point;
  * lengths are in units of 100 nm (WScale, vie_fft/ReadFile.h:7);
  * permittivities use the engineering sign exp(+j w t): loss is negative Im.
    Workload permittivities written in the physics form are conjugated once,
    here.
Random placement uses splitmix64 so any language regenerates the same media.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

WSCALE = 100.0          # nm per length unit (vie_fft/ReadFile.h:7)
EPS_BACKGROUND = 1.1    # dummy background permittivity (vie_fft/Model.h:82)
HC_EV_NM = 1239.8419843320026  # h c in eV nm


def phys_to_eng(eps):
    """Physics-form permittivity (loss Im > 0) to the reference's engineering form."""
    return np.conj(np.asarray(eps, dtype=np.complex128))


class SplitMix64:
    """splitmix64 with 53-bit uniform doubles in [0, 1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


@dataclass
class Sphere:
    center: tuple       # in cells
    radius: float       # in cells
    mat: int


@dataclass
class Band:
    z0: float           # [z0, z1) in cells, full extent in x and y
    z1: float
    mat: int


@dataclass
class Scene:
    n: tuple
    delta: float                      # cell size in length units (100 nm)
    wavelength_nm: float
    eps_table: np.ndarray             # engineering sign, index = material id
    shapes: list = field(default_factory=list)
    theta: float = 0.0
    phi: float = 0.0
    psi: float = 0.0
    area: float = 1.0                 # normalisation of Q_ext / Q_abs (length^2)
    absorb_mat: int = 1
    eps_voxel: np.ndarray = None      # optional per-voxel permittivity (design regions)

    @property
    def k0(self):
        return 2.0 * math.pi / (self.wavelength_nm / WSCALE)

    @property
    def d(self):
        return (self.delta, self.delta, self.delta)

    def material_at(self, x, y, z):
        """Material id at points given in cell units (arrays); 0 outside the box."""
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        z = np.asarray(z, dtype=np.float64)
        out = np.zeros(np.broadcast(x, y, z).shape, dtype=np.uint8)
        done = np.zeros(out.shape, dtype=bool)
        inside = (x >= 0) & (x <= self.n[0]) & (y >= 0) & (y <= self.n[1]) & (z >= 0) & (z <= self.n[2])
        for s in self.shapes:
            if isinstance(s, Sphere):
                cx, cy, cz = s.center
                hit = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= s.radius ** 2
            elif isinstance(s, Band):
                hit = (z >= s.z0) & (z < s.z1)
            else:
                raise TypeError(s)
            sel = hit & ~done & inside
            out[sel] = s.mat
            done |= sel
        return out

    def material_grid(self, xs, ys, zs):
        """material_at on the tensor grid xs x ys x zs (sorted 1-D coordinate
        arrays in cell units): each shape is evaluated on its bounding block only."""
        xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
        out = np.zeros((xs.size, ys.size, zs.size), dtype=np.uint8)
        done = np.zeros(out.shape, dtype=bool)

        def rng(a, lo, hi, nmax):  # index range of a within [lo, hi] and inside [0, nmax]
            lo, hi = max(lo, 0.0), min(hi, float(nmax))
            return np.searchsorted(a, lo, "left"), np.searchsorted(a, hi, "right")

        for s in self.shapes:
            if isinstance(s, Sphere):
                cx, cy, cz = s.center
                r = s.radius
                i0, i1 = rng(xs, cx - r, cx + r, self.n[0])
                j0, j1 = rng(ys, cy - r, cy + r, self.n[1])
                k0, k1 = rng(zs, cz - r, cz + r, self.n[2])
                if i0 >= i1 or j0 >= j1 or k0 >= k1:
                    continue
                hit = ((xs[i0:i1, None, None] - cx) ** 2 + (ys[None, j0:j1, None] - cy) ** 2
                       + (zs[None, None, k0:k1] - cz) ** 2) <= r * r
            elif isinstance(s, Band):
                i0, i1 = rng(xs, 0.0, self.n[0], self.n[0])
                j0, j1 = rng(ys, 0.0, self.n[1], self.n[1])
                k0 = np.searchsorted(zs, max(s.z0, 0.0), "left")
                k1 = np.searchsorted(zs, s.z1, "left")  # z < z1
                k1 = min(k1, np.searchsorted(zs, float(self.n[2]), "right"))
                if i0 >= i1 or j0 >= j1 or k0 >= k1:
                    continue
                hit = np.ones((i1 - i0, j1 - j0, k1 - k0), dtype=bool)
            else:
                raise TypeError(s)
            blk = (slice(i0, i1), slice(j0, j1), slice(k0, k1))
            sel = hit & ~done[blk]
            out[blk][sel] = s.mat
            done[blk] |= sel
        return out

    def material(self):
        return self.material_grid(*[np.arange(v, dtype=np.float64) + 0.5 for v in self.n])

    def flags(self):
        """Flag_X, Flag_Y, Flag_Z with the 3-point rule (vie_fft/Model.h:345-394):
        a face is active when the two cell centres it separates and the face
        centre are all non-background."""
        c = [np.arange(v, dtype=np.float64) + 0.5 for v in self.n]
        out = np.zeros((3,) + tuple(self.n), dtype=np.uint8)
        for f in range(3):
            m = []
            for shift in (0.5, 0.0, 1.0):   # face centre, lower cell, upper cell
                p = list(c)
                p[f] = c[f] + shift
                m.append(self.material_grid(*p) != 0)
            out[f] = (m[0] & m[1] & m[2]).astype(np.uint8)
        return out

    def eps(self):
        if self.eps_voxel is not None:
            return self.eps_voxel
        return self.eps_table[self.material()]

    def rhs_angles(self):
        return self.theta, self.phi, self.psi


def _table(*eps_eng):
    return np.array([EPS_BACKGROUND] + list(eps_eng), dtype=np.complex128)


def sphere_case(n=32, radius=10.0, eps_phys=-4.0 + 0.7j, delta_nm=2.0, wavelength_nm=400.0, psi=0.0):
    """Single sphere centred in an n^3 grid (the reference's default case has a
    60 nm Au sphere; this one takes the permittivity as an argument)."""
    c = n / 2.0
    sc = Scene(n=(n, n, n), delta=delta_nm / WSCALE, wavelength_nm=wavelength_nm,
               eps_table=_table(phys_to_eng(eps_phys)), shapes=[Sphere((c, c, c), radius, 1)], psi=psi)
    sc.area = math.pi * (radius * sc.delta) ** 2
    sc.absorb_mat = 1
    return sc


def layered_case(n=(128, 128, 64), band_start=8, sphere_z=36.0, delta_nm=2.0, wavelength_nm=550.0):
    """Layered slab with a 4x4 array of embedded spheres: bands [s, s+16) eps 2.1,
    [s+16, s+40) eps 3.6+0.35i, [s+40, s+48) eps -15+1i (physics form), spheres
    R = 6 cells eps -10+0.8i centred at ((16+32a), (16+32b), sphere_z)."""
    s = band_start
    table = _table(*phys_to_eng([2.1, 3.6 + 0.35j, -15.0 + 1.0j, -10.0 + 0.8j]))
    shapes = [Sphere((16 + 32 * a, 16 + 32 * b, sphere_z), 6.0, 4) for a in range(4) for b in range(4)]
    shapes += [Band(s, s + 16, 1), Band(s + 16, s + 40, 2), Band(s + 40, s + 48, 3)]
    sc = Scene(n=tuple(n), delta=delta_nm / WSCALE, wavelength_nm=wavelength_nm, eps_table=table, shapes=shapes)
    sc.area = float(n[0] * n[1]) * sc.delta ** 2
    sc.absorb_mat = 4
    return sc


def metric_case():
    """128^3 metric grid: the layered layout with the band stack from k = 40."""
    return layered_case(n=(128, 128, 128), band_start=40, sphere_z=68.0)


def random_spheres_case(n=256, count=60, seed=0x5EED0003, delta_nm=1.5, wavelength_nm=550.0):
    """Seeded random spheres: per sphere cx, cy, cz ~ U[0, n), R ~ U[5, 15], material
    index floor(3 U) into {-4+0.7i, -10+0.8i, 3.6+0.35i} (physics form)."""
    rng = SplitMix64(seed)
    table = _table(*phys_to_eng([-4.0 + 0.7j, -10.0 + 0.8j, 3.6 + 0.35j]))
    shapes = []
    for _ in range(count):
        cx, cy, cz = n * rng.uniform(), n * rng.uniform(), n * rng.uniform()
        r = 5.0 + 10.0 * rng.uniform()
        m = 1 + min(int(3 * rng.uniform()), 2)
        shapes.append(Sphere((cx, cy, cz), r, m))
    sc = Scene(n=(n, n, n), delta=delta_nm / WSCALE, wavelength_nm=wavelength_nm, eps_table=table, shapes=shapes)
    sc.area = float(n * n) * sc.delta ** 2
    return sc


def drude_eps_phys(wavelength_nm, eps_inf=9.5, wp_ev=8.95, gamma_ev=0.069):
    w = HC_EV_NM / np.asarray(wavelength_nm, dtype=np.float64)
    return eps_inf - wp_ev ** 2 / (w * (w + 1j * gamma_ev))


def dimer_case(wavelength_nm, psi=0.0, n=96, radius=12.0, delta_nm=1.0):
    """Au dimer along x (Drude), radius 12 cells, gap 4 cells, centres (34,48,48)
    and (62,48,48) in a 96^3 grid."""
    table = _table(phys_to_eng(drude_eps_phys(wavelength_nm)))
    shapes = [Sphere((34.0, 48.0, 48.0), radius, 1), Sphere((62.0, 48.0, 48.0), radius, 1)]
    sc = Scene(n=(n, n, n), delta=delta_nm / WSCALE, wavelength_nm=wavelength_nm, eps_table=table,
               shapes=shapes, psi=psi)
    sc.area = 2 * math.pi * (radius * sc.delta) ** 2
    return sc


def dimer_sweep(count=64):
    lams = 400.0 + 400.0 * np.arange(count) / (count - 1)
    return [(lam, psi) for lam in lams for psi in (0.0, 90.0)]


def design_case(n=64, seed=0x5EED0005, delta_nm=5.0, wavelength_nm=600.0):
    """Whole grid is a design region: eps_v = U[1, 12] + 0.1i (physics form)."""
    rng = SplitMix64(seed)
    V = n ** 3
    re = np.array([1.0 + 11.0 * rng.uniform() for _ in range(V)]).reshape(n, n, n)
    eps = phys_to_eng(re + 0.1j)
    shapes = [Band(-1.0, n + 1.0, 1)]  # every in-box point is non-background
    sc = Scene(n=(n, n, n), delta=delta_nm / WSCALE, wavelength_nm=wavelength_nm,
               eps_table=_table(phys_to_eng(6.5 + 0.1j)), shapes=shapes)
    sc.eps_voxel = eps
    sc.area = float(n * n) * sc.delta ** 2
    return sc
