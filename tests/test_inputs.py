"""Benchmark inputs against the reference's own forcing test values
(physics/test/BenchmarkBoundaries_test.cpp:20-82)."""
import numpy as np

from nextsimdg_amd import inputs
from oracle import Oracle


def test_ocean_gyre_corner_values():
    nx = ny = 256
    u, v = inputs.ocean_gyre(nx, ny)
    u, v = u.reshape(ny, nx), v.reshape(ny, nx)
    vmax = 0.01
    assert u[0, 0] == -vmax and v[0, 0] == vmax
    assert u[ny - 1, nx - 1] == (ny - 2.0) / ny * vmax
    assert v[ny - 1, nx - 1] == -(nx - 2.0) / nx * vmax
    assert u[40, 50] != 0 and v[40, 50] != 0


def test_cyclone_moves_away_from_test_point():
    """The wind at (i, j) = (50, 40) weakens over the first hour (the cyclone recedes)."""
    nx = ny = 256
    ora = Oracle()
    u0, v0 = ora.benchmark_wind(nx, ny, inputs.L_BOX, 0.0)
    u1, v1 = ora.benchmark_wind(nx, ny, inputs.L_BOX, 3600.0)
    e = 50 + nx * 40
    assert u0[e] != 0 and v0[e] != 0
    assert abs(u1[e]) < abs(u0[e]) and abs(v1[e]) < abs(v0[e])
    # peak speed of the reference cyclone: 30 m/s * max(r k exp(-k r)) = 30/e at r = 1/k
    speed = np.hypot(u0, v0)
    assert speed.max() < 30.0 / np.e + 1e-9
    assert speed.max() > 0.95 * 30.0 / np.e


def test_box_mask_and_coords():
    c, m = inputs.box_coords(8, 6), inputs.box_mask(8, 6).reshape(6, 8)
    assert c.shape == (63, 2) and c[-1, 0] == inputs.L_BOX and c[-1, 1] == inputs.L_BOX
    assert m[0].sum() == 0 and m[:, 0].sum() == 0 and m[1:-1, 1:-1].min() == 1
