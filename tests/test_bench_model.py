"""The bench's algorithmic byte model (CPU only): the dense BiCGSTAB iteration
bytes reproduce the per-configuration figures of the path's measurement plan
(B_iter = 2 (81 V s + octant s + 4 V) + 48 V s), the windowed model charges
only the window's cells, and the x-pass instantiation follows launch_x."""
import pytest

import bench

# grid, FP64 and FP32 bytes per BiCGSTAB iteration (GB) from the plan's table
PLAN_TABLE = [((32, 32, 32), 0.1115, 0.0559), ((128, 128, 64), 3.566, 1.787), ((128, 128, 128), 7.132, 3.574),
              ((256, 256, 256), 57.05, 28.59), ((96, 96, 96), 3.009, 1.508), ((64, 64, 64), 0.892, 0.447)]


def full_plan(n):
    return {"axis_order": (0, 1, 2), "x_window": (0, n[0]), "x_pencil": n[0], "grid": tuple(n),
            "y_window": (0, n[1]), "y_pencil": n[1]}


@pytest.mark.parametrize("n,gb64,gb32", PLAN_TABLE)
def test_dense_iteration_bytes_match_the_plan(n, gb64, gb32):
    d64, w64 = bench.solve_models(full_plan(n), 16.0, 1)
    d32, w32 = bench.solve_models(full_plan(n), 8.0, 1)
    assert abs(d64 / 1e9 - gb64) <= 0.5e-3 * gb64 + 1e-4
    assert abs(d32 / 1e9 - gb32) <= 0.5e-3 * gb32 + 1e-4
    # a window covering the grid charges what the dense model does
    assert w64 == pytest.approx(d64) and w32 == pytest.approx(d32)


def test_window_model_scales_with_the_window():
    plan = dict(full_plan((128, 128, 128)), x_window=(40, 88), x_pencil=64)
    d, w = bench.solve_models(plan, 8.0, 10)
    V, vw = 128 ** 3, 48 * 128 * 128
    per_it = 2 * (81.0 * vw * 8 + 65 * 129 * 129 * 8 + 4.0 * vw) + 48.0 * vw * 8
    assert w == pytest.approx(10 * per_it)
    assert d == pytest.approx(10 * 3.5743396e9, rel=1e-6)
    # a mixed solve adds one FP64 product and 12 V FP64 entries per replacement
    _, w2 = bench.solve_models(plan, 8.0, 10, outer=2)
    assert w2 - w == pytest.approx(2 * (bench.product_bytes(plan, 16.0) + 12.0 * V * 16))


def test_xpass_bytes_of_the_workload():
    plan = dict(full_plan((128, 128, 128)), x_window=(40, 88), x_pencil=64)
    assert bench.xpass_bytes(plan, 8.0) == 2 * 3 * 256 * 256 * 48 * 8 + 65 * 129 * 129 * 8 == 159648264


@pytest.mark.parametrize("pencil,window,grid_x,sup", [(64, (40, 88), 128, 48), (64, (10, 40), 128, 32),
                                                      (64, (0, 60), 128, 64), (128, (0, 90), 128, 96),
                                                      (128, (0, 128), 128, 128), (32, (5, 25), 64, 24)])
def test_x_pass_instantiation(pencil, window, grid_x, sup):
    plan = dict(full_plan((grid_x, 16, 16)), x_window=window, x_pencil=pencil)
    assert bench.xreg_support(plan) == sup
