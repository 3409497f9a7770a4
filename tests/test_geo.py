"""Map output conventions (ext/AutoRIFTRastersExt.jl, ext/AutoRIFTDimensionalDataExt.jl): signs,
north-up / south-up storage, dt spellings, grid coordinates. CPU only: results are built by hand.
The end-to-end run on the GPU (velocities from a tracked pair) is in test_gpu_app.py."""
import datetime

import numpy as np
import pytest

import autorift
from autorift.core import MultichipResult


def _result(dx, dy):
    dx = np.asarray(dx, np.float32)
    dy = np.asarray(dy, np.float32)
    z = np.zeros(dx.shape)
    return MultichipResult(dx, dy, z.astype(np.float32), z.astype(np.uint16), z.astype(bool))


RES = 30.0
NORTH_UP = 1000.0 - RES * np.arange(50)      # y decreasing
SOUTH_UP = NORTH_UP[::-1].copy()
X = 500.0 + RES * np.arange(60)


def test_signs_north_up():
    """vx = -dx; on a north-up raster vy = +dy (the row-down and north-up flips cancel),
    test/extensions.jl:114-121."""
    r = _result([[-6.0, 2.5]], [[4.0, -1.0]])
    vx, vy = autorift.to_velocity(r, NORTH_UP)
    np.testing.assert_array_equal(vx, -r.dx)
    np.testing.assert_array_equal(vy, r.dy)
    assert vx.dtype == np.float32 and vy.dtype == np.float32


def test_storage_order_gives_the_same_field():
    """A south-up copy of a scene gives dy of the opposite sign and reversed rows; the velocity
    field is identical (test/extensions.jl:48-94)."""
    rng = np.random.default_rng(0)
    dx = rng.normal(size=(7, 9)).astype(np.float32)
    dy = rng.normal(size=(7, 9)).astype(np.float32)
    north = autorift.to_velocity(_result(dx, dy), NORTH_UP)
    south = autorift.to_velocity(_result(dx[::-1], -dy[::-1]), SOUTH_UP)
    for a, b in zip(north, south):
        np.testing.assert_array_equal(a, b[::-1])
    # features moving 6 px east and 4 px south: core offsets (-6, -4) on a north-up raster
    vx, vy = autorift.to_velocity(_result([[-6.0]], [[-4.0]]), NORTH_UP)
    assert (vx[0, 0], vy[0, 0]) == (6.0, -4.0)


def test_dt_spellings():
    """Day counts, date differences and years agree; calendar periods, non-positive intervals
    and other types are refused (test/extensions.jl:188-235)."""
    r = _result([[-6.0]], [[-4.0]])
    yrs = 16 * 86400e3 / autorift.geo.MS_PER_YEAR
    vx, vy = autorift.to_velocity(r, NORTH_UP, dt=datetime.timedelta(days=16), x=X)
    assert vx[0, 0] == pytest.approx(6 * RES / yrs, rel=1e-6)
    assert vy[0, 0] == pytest.approx(-4 * RES / yrs, rel=1e-6)
    d = np.datetime64("2020-07-01") - np.datetime64("2020-06-15")
    vx2, _ = autorift.to_velocity(r, NORTH_UP, dt=d, x=X)
    assert vx2[0, 0] == pytest.approx(6 * RES / yrs, rel=1e-6)
    vx3, _ = autorift.to_velocity(r, NORTH_UP, dt=1.0, x=X)
    assert vx3[0, 0] == pytest.approx(6 * RES, rel=1e-6)
    assert autorift.years(np.timedelta64(16, "D")) == pytest.approx(yrs)
    with pytest.raises(autorift.ArgumentError, match="Day"):
        autorift.years(np.timedelta64(1, "Y"))
    with pytest.raises(autorift.ArgumentError):
        autorift.years(np.timedelta64(6, "M"))
    for bad in (-1.0, 0, "16 days", datetime.timedelta(days=-1)):
        with pytest.raises(autorift.ArgumentError):
            autorift.years(bad)


def test_irregular_axis_refuses_dt_only():
    irr = NORTH_UP.copy()
    irr[10] += 3.0
    r = _result([[-6.0]], [[0.0]])
    with pytest.raises(autorift.ArgumentError, match="regularly spaced"):
        autorift.to_velocity(r, irr, dt=datetime.timedelta(days=16), x=X)
    vx, _ = autorift.to_velocity(r, irr)           # no dt: no pixel size needed
    assert vx[0, 0] == 6.0
    assert autorift.pixel_size(X) == RES and autorift.pixel_size(irr) is None


def test_grid_coordinates():
    """Output coordinates: the input's, sampled at the grid's 1-based positions with the grid's
    regular step (ext/AutoRIFTDimensionalDataExt.jl:113-144)."""
    grid = autorift.gridpoints((50, 60), 8, chip_size=4, search_radius=3)
    gy, gx = autorift.grid_coordinates(NORTH_UP, X, grid)
    rows = grid.y[:, 0].astype(int)
    cols = grid.x[0, :].astype(int)
    np.testing.assert_array_equal(gy, NORTH_UP[rows - 1])
    np.testing.assert_array_equal(gx, X[cols - 1])
    assert autorift.pixel_size(gy) == 8 * RES and gy[1] < gy[0]
