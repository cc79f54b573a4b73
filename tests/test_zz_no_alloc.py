"""vie_far_field_noalloc (static device scratch, no allocation) against
vie_far_field on the same resident solution: two calls in a row, the second
with fewer directions, and the size limit."""
import numpy as np
import pytest

from vie_fft_amd import geometry as G
from vie_fft_amd import viefft as VF

pytestmark = pytest.mark.gpu


def test_far_field_noalloc_equals_far_field():
    sc = G.sphere_case(24, radius=7.0, eps_phys=2.25 + 0.2j, wavelength_nm=500.0)
    with VF.Solver(sc.n, sc.d, sc.k0) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.set_rhs_incident(0.0, 0.0, 0.0)
        s.solve_resident(VF.BICGSTAB, tol=1e-8)
        rng = np.random.default_rng(3)
        for ndir in (512, 37):   # the second call is smaller: no stale partials may leak in
            th = rng.uniform(0, 180, ndir)
            ph = rng.uniform(0, 360, ndir)
            ref = s.far_field(th, ph)
            got = s.far_field_noalloc(th, ph)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"ndir {ndir}: max rel diff {err:.2e}", flush=True)
            assert err <= 1e-12
        with pytest.raises(VF.VieError) as ei:
            s.far_field_noalloc(np.zeros(513), np.zeros(513))
        assert ei.value.status == VF.ERR_ARG
        assert s.far_field_noalloc(np.zeros(0), np.zeros(0)).shape == (0, 3)
