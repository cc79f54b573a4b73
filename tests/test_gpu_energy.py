"""Energy balance and reciprocity of GPU solutions through the C ABI.

Energy balance: vie_forward_scatter's
extinction equals vie_absorption's absorbed power plus the scattered power
integrated over the sphere from vie_far_field's amplitudes (the relation and
its quadrature are those of tests/test_oracle_energy.py, which pins them on
the oracle).  Media: the C1 plasmonic sphere (32^3, R = 10 cells,
eps -4+0.7i, 400 nm), a lossless dielectric sphere, and a small layered slab
with embedded spheres; FP64 solves to 1e-11 and a mixed solve to 1e-8.

Reciprocity (the relation of tests/test_oracle_reciprocity.py): the device-built
plane waves (vie_set_rhs_incident) from two oblique directions with rotated
polarisations, and vie_far_field's amplitudes towards the opposite directions."""
import numpy as np
import pytest

from tests.gpu_util import scene_solver
from tests.test_oracle_energy import Z0, q_sca, sphere_directions
from tests.test_oracle_reciprocity import A1, A2, backward, polarisation
from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu

CASES = {
    "c1_plasmonic_sphere": lambda: G.sphere_case(32, radius=10.0, eps_phys=-4.0 + 0.7j, delta_nm=2.0,
                                                 wavelength_nm=400.0),
    "lossless_sphere": lambda: G.sphere_case(24, radius=8.0, eps_phys=4.0 + 0.0j, delta_nm=4.0, wavelength_nm=500.0),
    "layered_slab": lambda: G.layered_case(n=(32, 32, 52), band_start=2, sphere_z=30.0),
}


def _balance(s, sc):
    th, ph, wt = sphere_directions()
    S = s.far_field(np.degrees(th), np.degrees(ph))
    qs = q_sca(S, th, ph, wt, sc.k0, sc.area)
    qe = s.forward_scatter(*sc.rhs_angles()) / sc.area
    qa = s.absorption() * Z0 / sc.area
    return qe, qa, qs


@pytest.mark.parametrize("name", sorted(CASES))
def test_extinction_equals_absorption_plus_scattering(name):
    sc = CASES[name]()
    with scene_solver(sc) as s:
        s.set_rhs_incident(*sc.rhs_angles())
        info = s.solve_resident(VF.BICGSTAB, tol=1e-11, maxit=20000)
        assert info["converged"]
        qe, qa, qs = _balance(s, sc)
        print(f"{name}: Q_ext {qe:.8f} Q_abs {qa:.8f} Q_sca {qs:.8f} balance {(qa + qs) / qe - 1:.2e}")
        assert qe > 0 and qs > 0
        assert abs(qa + qs - qe) <= 1e-4 * qe, (qe, qa, qs)
        if name == "c1_plasmonic_sphere":
            # the mixed solve's outputs keep the balance as well
            mixed = s.solve_mixed(tol=1e-8)
            assert mixed["converged"]
            qe2, qa2, qs2 = _balance(s, sc)
            assert abs(qa2 + qs2 - qe2) <= 1e-4 * qe2, (qe2, qa2, qs2)
            assert abs(qe2 - qe) <= 1e-6 * qe


@pytest.mark.parametrize("name", ["c1_plasmonic_sphere", "layered_slab"])
def test_far_fields_are_reciprocal(name):
    sc = CASES[name]()
    with scene_solver(sc) as s:
        amp = []
        for a, other in ((A1, A2), (A2, A1)):
            s.set_rhs_incident(*a)
            assert s.solve_resident(VF.BICGSTAB, tol=1e-11, maxit=20000)["converged"]
            t, p = backward(*other[:2])
            amp.append(polarisation(*other) @ s.far_field([np.degrees(t)], [np.degrees(p)])[0])
    u, v = amp
    print(f"{name}: reciprocity {abs(u - v) / abs(u):.2e}")
    assert abs(u - v) <= 1e-4 * abs(u), (u, v)
