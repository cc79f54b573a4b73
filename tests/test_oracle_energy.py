"""Energy balance of the oracle's VIE solutions: the extinction that
FarField1's forward sum gives (optical theorem, vie_fft/NearFar.h:16-109)
equals the absorbed power of NearField's face sum (vie_fft/NearFar.h:180-250)
plus the scattered power, integrated here over the whole sphere of
directions from the same far-field sums (FarField2's, vie_fft/NearFar.h:302-456):

    Q_ext = Z0 Re(N(k_inc) . e) / A,   Q_abs = Z0 P_abs / A,
    Q_sca = k0^2 Z0^2 / (16 pi^2 A) * integral of |N_theta|^2 + |N_phi|^2 over the sphere,

N the far-field sums dV sum_c J_c e^{j k0 r.v} at face centres, A the
normalisation area.  Poynting's theorem makes the balance exact for the
continuous problem; the discrete one keeps it to the far-field sums'
point-current error, (k0 delta)^2 small here.  A dropped factor (2, pi, Z0),
a wrong sign or a swapped component in any of the three outputs breaks it by
far more than the 1e-4 this test allows."""
import numpy as np
import pytest

from oracle import oracle as O
from vie_fft_amd import geometry as G

Z0 = 120 * np.pi


def far_sums(n, d, k0, J, th, ph):
    """The oracle's far-field sums at many directions (th, ph in radians),
    restated with separable phases."""
    J = J.reshape((3,) + tuple(n))
    pos = [(np.arange(n[a]) + 0.5) * d[a] for a in range(3)]
    vd = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    e = [np.exp(1j * k0 * np.outer(vd[:, a], pos[a])) for a in range(3)]
    S = np.empty((len(th), 3), complex)
    for c in range(3):
        t = np.einsum("ijk,dk->dij", J[c], e[2])
        t = np.einsum("dij,dj->di", t, e[1])
        S[:, c] = np.einsum("di,di->d", t, e[0]) * np.exp(0.5j * k0 * d[c] * vd[:, c]) * np.prod(d)
    return S


def sphere_directions(nt=48, nph=96):
    """Gauss-Legendre in cos(theta) times a uniform phi rule: (theta, phi, weight)."""
    x, w = np.polynomial.legendre.leggauss(nt)
    th, ph = np.meshgrid(np.arccos(x), 2 * np.pi * np.arange(nph) / nph, indexing="ij")
    wt = np.outer(w, np.full(nph, 2 * np.pi / nph))
    return th.ravel(), ph.ravel(), wt.ravel()


def q_sca(S, th, ph, wt, k0, area):
    Nt = S[:, 0] * np.cos(th) * np.cos(ph) + S[:, 1] * np.cos(th) * np.sin(ph) - S[:, 2] * np.sin(th)
    Np = -S[:, 0] * np.sin(ph) + S[:, 1] * np.cos(ph)
    return k0 ** 2 * Z0 ** 2 / (16 * np.pi ** 2 * area) * np.sum((abs(Nt) ** 2 + abs(Np) ** 2) * wt)


def test_rule_integrates_the_sphere():
    th, ph, wt = sphere_directions()
    assert abs(wt.sum() - 4 * np.pi) < 1e-12
    # a dipole pattern sin^2(theta) integrates to 8 pi / 3
    assert abs(np.sum(np.sin(th) ** 2 * wt) - 8 * np.pi / 3) < 1e-12


def test_far_sums_restatement_matches_forward_scatter():
    sc = G.sphere_case(8, radius=3.0, eps_phys=3.0 + 0.2j, delta_nm=4.0)
    rng = np.random.default_rng(3)
    J = (rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)) * sc.flags()
    S = far_sums(sc.n, sc.d, sc.k0, J, np.array([0.0]), np.array([0.0]))[0]
    assert abs(Z0 * S[0].real - O.forward_scatter(sc.n, sc.d, sc.k0, J)) < 1e-12 * abs(Z0 * S[0].real)


@pytest.mark.parametrize("eps,delta_nm", [(4.0 + 0.0j, 4.0), (4.0 + 0.5j, 4.0), (-4.0 + 0.7j, 2.0)])
def test_extinction_equals_absorption_plus_scattering(eps, delta_nm):
    sc = G.sphere_case(16, radius=6.0, eps_phys=eps, delta_nm=delta_nm, wavelength_nm=400.0)
    n, d, k0 = sc.n, sc.d, sc.k0
    e, flags = sc.eps(), sc.flags()
    op = O.Operator(n, d, k0, e, flags, O.SCHED_FUSED)
    b = O.incident(n, d, k0, *sc.rhs_angles()) * flags
    r = O.solve(3, lambda x, adj: op.apply(x.reshape((3,) + tuple(n)), adjoint=adj).ravel(), b.ravel(),
                tol=1e-11, maxit=5000)
    assert r["converged"]
    J = r["x"].reshape((3,) + tuple(n))
    qe = O.forward_scatter(n, d, k0, J, *sc.rhs_angles()) / sc.area
    qa = O.absorption(n, d, k0, J, e, flags) * Z0 / sc.area
    th, ph, wt = sphere_directions()
    qs = q_sca(far_sums(n, d, k0, J, th, ph), th, ph, wt, k0, sc.area)
    if eps.imag == 0:
        assert abs(qa) < 1e-12  # lossless: nothing absorbed
    assert qs > 0 and qe > 0
    assert abs(qa + qs - qe) <= 1e-4 * qe, (qe, qa, qs)
