"""The block-evaluated material and flag builders equal the pointwise
3-point rule (vie_fft/Model.h:345-394) evaluated at every point."""
import numpy as np
import pytest

from vie_fft_amd import geometry as G


def _pointwise(sc):
    i, j, k = np.meshgrid(*[np.arange(v, dtype=np.float64) for v in sc.n], indexing="ij")
    c = [i + 0.5, j + 0.5, k + 0.5]
    mat = sc.material_at(*c)
    flags = np.zeros((3,) + tuple(sc.n), np.uint8)
    for f in range(3):
        p = [a.copy() for a in c]
        p[f] = c[f] + 0.5
        m1 = sc.material_at(*p)
        p[f] = c[f]
        m2 = sc.material_at(*p)
        p[f] = c[f] + 1.0
        m3 = sc.material_at(*p)
        flags[f] = (m1 != 0) & (m2 != 0) & (m3 != 0)
    return mat, flags


@pytest.mark.parametrize("make", [
    lambda: G.sphere_case(n=24, radius=7.5),
    lambda: G.layered_case(n=(40, 36, 32), band_start=6, sphere_z=18.0),
    lambda: G.random_spheres_case(n=40, count=25),
    lambda: G.dimer_case(520.0, n=40, radius=5.0),
    lambda: G.design_case(n=8),
])
def test_block_geometry_equals_pointwise(make):
    sc = make()
    mat, flags = _pointwise(sc)
    assert np.array_equal(sc.material(), mat)
    assert np.array_equal(sc.flags(), flags)
    # the own-axis last plane is never active (its upper cell is outside)
    assert not flags[0, -1].any() and not flags[1, :, -1].any() and not flags[2, :, :, -1].any()
