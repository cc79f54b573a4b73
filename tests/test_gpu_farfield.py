"""GPU far-field sums and FarField2 RCS cuts against the oracle
(vie_fft/NearFar.h:16-109, 302-456)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import scene_solver
from vie_fft_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solved():
    sc = G.layered_case(n=(24, 20, 16), band_start=2, sphere_z=8.0)
    s = scene_solver(sc)
    s.set_rhs_incident(0.0, 0.0, 30.0)
    s.solve_resident(tol=1e-10)
    J = s.solution()
    yield sc, s, J
    s.close()


@pytest.mark.parametrize("phi", [0.0, 45.0])
def test_rcs_matches_oracle(solved, phi):
    sc, s, J = solved
    rt, rp = s.rcs(phi)
    ot, op = O.rcs(sc.n, sc.d, sc.k0, J, 0.0, phi)
    for got, want in ((rt, ot), (rp, op)):
        ok = np.isfinite(want) & (want > np.nanmax(want) - 100)
        assert ok.sum() > 150
        assert np.max(np.abs(got[ok] - want[ok])) < 1e-9


def test_far_field_reproduces_forward_scatter(solved):
    """FarField1's extinction from the multi-direction amplitudes equals vie_forward_scatter."""
    sc, s, J = solved
    Z0 = 120 * np.pi
    for th, ph, ps in ((0.0, 0.0, 0.0), (0.0, 0.0, 30.0), (25.0, 40.0, 10.0)):
        S = s.far_field([th], [ph])[0]
        t, p, q = np.radians([th, ph, ps])
        tm = Z0 * (S[0].real * np.cos(t) * np.cos(p) + S[1].real * np.cos(t) * np.sin(p) - S[2].real * np.sin(t))
        te = Z0 * (S[0].real * np.sin(p) - S[1].real * np.cos(p))
        fs = s.forward_scatter(th, ph, ps)
        assert abs(np.cos(q) * tm - np.sin(q) * te - fs) <= 1e-11 * abs(fs)
        assert abs(fs - O.forward_scatter(sc.n, sc.d, sc.k0, J, th, ph, ps)) <= 1e-11 * abs(fs)


def test_far_field_many_directions(solved):
    """More directions than one launch holds (chunks of 1024) give the same values."""
    sc, s, J = solved
    rng = np.random.default_rng(4)
    th = rng.uniform(0, 180, 2500)
    ph = rng.uniform(0, 360, 2500)
    S = s.far_field(th, ph)
    S2 = np.concatenate([s.far_field(th[i:i + 700], ph[i:i + 700]) for i in range(0, 2500, 700)])
    assert np.array_equal(S, S2)
    with pytest.raises(ValueError):
        s.far_field([0.0, 1.0], [0.0])
