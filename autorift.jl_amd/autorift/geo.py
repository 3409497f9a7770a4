"""Map output conventions: grid coordinates, north-up / south-up handling and pixel offsets to
velocity, as AutoRIFT.jl's Rasters and DimensionalData extensions do
(ext/AutoRIFTRastersExt.jl:89-170, ext/AutoRIFTDimensionalDataExt.jl:84-182).

There is no raster type here: a caller passes the reference's pixel-centre coordinates along
rows (`y`) and columns (`x`), in CRS units, with the images as plain arrays. The core reports
the offset from secondary back to reference in array rows and columns; `to_velocity` turns it
into feature motion in map orientation (`+vx` east, `+vy` north), optionally per year.
"""
from __future__ import annotations

import datetime

import numpy as np

from ._lib import ArgumentError, DimensionMismatch

# Milliseconds in a Julian year, the convention ITS_LIVE publishes velocities in.
MS_PER_YEAR = 365.25 * 24 * 60 * 60 * 1000


def pixel_size(coords) -> float | None:
    """|step| of a regularly spaced axis, or None (ext/AutoRIFTDimensionalDataExt.jl:182)."""
    c = np.asarray(coords, dtype=np.float64)
    if c.size < 2:
        return None
    d = np.diff(c)
    if not np.allclose(d, d[0], rtol=1e-9, atol=0.0) or d[0] == 0:
        return None
    return float(abs(d[0]))


def sample_lookup(coords, positions) -> np.ndarray:
    """Coordinates at 1-based pixel positions, stepping regularly from the first to the last
    (ext/AutoRIFTDimensionalDataExt.jl:138-144)."""
    c = np.asarray(coords)
    pos = np.asarray(positions, dtype=np.float64).ravel()
    first = int(round(pos[0]))
    if pos.size == 1:
        return c[first - 1:first]
    step = int(round(pos[1] - pos[0]))
    last = int(round(pos[-1]))
    return c[first - 1:last:step]


def grid_coordinates(y, x, grid):
    """(y, x) coordinates of a gridded point set's rows and columns
    (ext/AutoRIFTDimensionalDataExt.jl:113-121): the grid's y varies down rows, x across
    columns."""
    gy = np.asarray(grid.y)
    gx = np.asarray(grid.x)
    if gy.ndim != 2:
        raise ArgumentError("grid coordinates need a gridded (2-D) point set")
    return sample_lookup(y, gy[:, 0]), sample_lookup(x, gx[0, :])


def years(dt) -> float:
    """The acquisition interval in Julian years (ext/AutoRIFTRastersExt.jl:155-168). A number is
    years; a datetime.timedelta or a fixed-unit numpy.timedelta64 converts exactly; calendar
    units (years, months) have no fixed length and are refused."""
    if isinstance(dt, (bool, np.bool_)):
        raise ArgumentError(f"`dt` must be a time interval or a number of years, got {dt!r}")
    if isinstance(dt, datetime.timedelta):
        ms = dt / datetime.timedelta(milliseconds=1)
        if ms <= 0:
            raise ArgumentError(f"`dt = {dt}` must be positive; it is the interval between the "
                                "two acquisitions.")
        return ms / MS_PER_YEAR
    if isinstance(dt, np.timedelta64):
        unit = np.datetime_data(dt.dtype)[0]
        if unit in ("Y", "M"):
            raise ArgumentError(
                f"`dt = {dt}` is a calendar period with no fixed length, so it cannot be "
                "converted to a velocity denominator. Use a fixed period such as "
                "numpy.timedelta64(365, 'D') (a Day count), a number of years, or the "
                "difference of two dates.")
        ms = dt / np.timedelta64(1, "ms")
        if not ms > 0:
            raise ArgumentError(f"`dt = {dt}` must be positive; it is the interval between the "
                                "two acquisitions.")
        return float(ms) / MS_PER_YEAR
    if isinstance(dt, (int, float, np.integer, np.floating)):
        if not dt > 0:
            raise ArgumentError(f"`dt = {dt}` must be positive; it is the interval between the "
                                "two acquisitions, in years.")
        return float(dt)
    raise ArgumentError("`dt` must be a datetime.timedelta, a numpy.timedelta64, a number of "
                        f"years, or None, got a {type(dt).__name__}.")


def to_velocity(result, y, dt=None, x=None):
    """(vx, vy) feature motion in map orientation from a MultichipResult
    (ext/AutoRIFTRastersExt.jl:140-150).

    `y` is the reference's row coordinates (decreasing on a north-up raster). vx = -dx always;
    vy = +dy on a north-up raster and -dy on a south-up one, so both storage orders of a scene
    give the same field. With `dt`, both are scaled to CRS units per year by the pixel sizes of
    `x` and `y` (which must then be regularly spaced)."""
    yc = np.asarray(y, dtype=np.float64)
    sy = np.float32(1.0) if yc.size >= 2 and yc[1] < yc[0] else np.float32(-1.0)
    dx = np.asarray(result.dx, dtype=np.float32)
    dy = np.asarray(result.dy, dtype=np.float32)
    if dt is None:
        return -dx, sy * dy
    yrs = years(dt)
    if x is None:
        raise ArgumentError("`dt` needs the column coordinates `x` for the pixel size")
    px, py = pixel_size(x), pixel_size(yc)
    if px is None or py is None:
        raise ArgumentError(
            "`dt` was given but the coordinates are not regularly spaced, so there is no single "
            "pixel size to scale by and the velocity would be wrong wherever the spacing "
            "differs. Resample onto a regular grid, or pass `dt = None` and convert the pixel "
            "displacements yourself.")
    return ((-dx.astype(np.float64) * (px / yrs)).astype(np.float32),
            (sy * dy.astype(np.float64) * (py / yrs)).astype(np.float32))


def autorift_geo(reference, secondary, y, x, dt=None, crs=None, secondary_crs=None,
                 reference_valid=None, secondary_valid=None, device=0, **kwargs):
    """autorift on a georeferenced pair (ext/AutoRIFTRastersExt.jl:89-111): the images share the
    pixel-centre coordinates `y` (rows) and `x` (columns) in one CRS. Returns the layers vx, vy,
    correlation, chip_size, interpolated on the output grid, with its coordinates."""
    from .api import autorift_with_grid
    ny, nx = np.shape(reference)
    if np.shape(secondary) != (ny, nx):
        raise DimensionMismatch(f"reference is {(ny, nx)} but secondary is "
                                f"{np.shape(secondary)}; the two images must be on a common grid")
    if len(y) != ny or len(x) != nx:
        raise DimensionMismatch(f"{len(y)} row and {len(x)} column coordinates for a {ny}x{nx} "
                                "image")
    if crs is not None and secondary_crs is not None and crs != secondary_crs:
        raise ArgumentError(f"the two images are in different coordinate systems, {crs} and "
                            f"{secondary_crs}, so they are not co-registered. Reproject one onto "
                            "the other's CRS before correlating.")
    result, grid = autorift_with_grid(reference, secondary, reference_valid=reference_valid,
                                      secondary_valid=secondary_valid, device=device, **kwargs)
    gy, gx = grid_coordinates(y, x, grid)
    # offsets are in input pixels: the input's orientation and pixel size convert them
    vx, vy = to_velocity(result, y, dt, x)
    return {"vx": vx, "vy": vy, "correlation": result.correlation,
            "chip_size": result.chip_size, "interpolated": np.asarray(result.interpolated, bool),
            "y": gy, "x": gx, "crs": crs,
            "missingval": {"vx": np.float32(np.nan), "vy": np.float32(np.nan),
                           "correlation": np.float32(np.nan), "chip_size": np.uint16(0),
                           "interpolated": False}}
