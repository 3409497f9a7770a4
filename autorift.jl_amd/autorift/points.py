"""PointSet and its constructors (src/points.jl): where to correlate and with what geometry.

Fields are numpy arrays of one shape: a 1-D shape is a scattered set (PointSet{1}), a 2-D shape a
grid (PointSet{2}). Coordinates are 1-based pixels, as in the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, fields, replace

import numpy as np

from ._lib import ArgumentError, DimensionMismatch


@dataclass
class PointSet:
    x: np.ndarray
    y: np.ndarray
    radius_x: np.ndarray
    radius_y: np.ndarray
    dx_prior: np.ndarray
    dy_prior: np.ndarray
    chip_size_x: np.ndarray
    chip_size_y: np.ndarray

    def __post_init__(self):
        shp = np.shape(self.x)
        for f in fields(self):
            if np.shape(getattr(self, f.name)) != shp:
                raise DimensionMismatch(
                    f"PointSet field `{f.name}` has shape {np.shape(getattr(self, f.name))}, "
                    f"expected {shp} to match `x`")
        # column-major storage, as the reference's arrays and the C ABI have it: handing the
        # fields to the library is then a view, not a transposing copy per call
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray) and v.ndim > 1 and not v.flags.f_contiguous:
                setattr(self, f.name, np.asfortranarray(v))

    @property
    def shape(self):
        return np.shape(self.x)

    def __len__(self):
        return int(np.prod(self.shape))

    def nsearchable(self) -> int:
        return int(np.count_nonzero((self.radius_x > 0) & (self.radius_y > 0)))

    def rebuild(self, **kw) -> "PointSet":
        """A copy with the named fields replaced (src/points.jl:309-313)."""
        return replace(self, **kw)

    def __getitem__(self, idx):
        """Decimate or crop a gridded set, copying every field (src/points.jl:293-297)."""
        return PointSet(*[np.array(getattr(self, f.name)[idx]) for f in fields(self)])


def _field(v, like, dtype, name):
    if np.isscalar(v):
        if np.issubdtype(dtype, np.integer) and float(v) != int(v):
            raise ArgumentError(f"expected an integer value, got {v}")
        return np.full(np.shape(like), v, dtype=dtype)
    a = np.asarray(v)
    if a.shape != np.shape(like):
        raise DimensionMismatch(f"`{name}` has shape {a.shape} but there are "
                                f"{np.shape(like)} points")
    if np.issubdtype(dtype, np.integer) and not np.all(np.asarray(a, float) == np.round(a)):
        raise ArgumentError(f"`{name}` must hold integers")
    return a.astype(dtype)


def _round_even(v):
    return 2 * np.rint(np.asarray(v, float) / 2).astype(np.int64)


def pointset(x, y, search_radius=25, search_radius_x=None, search_radius_y=None, chip_size=32,
             chip_size_x=None, chip_size_y=None, chip_aspect=1.0, dx_prior=0.0, dy_prior=0.0):
    """src/points.jl:149-182."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if x.shape != y.shape:
        raise DimensionMismatch(f"`x` and `y` must have the same shape, got {x.shape} and {y.shape}")
    rx0, ry0 = search_radius if isinstance(search_radius, tuple) else (search_radius, search_radius)
    rx = _field(rx0 if search_radius_x is None else search_radius_x, x, np.int64, "search_radius_x")
    ry = _field(ry0 if search_radius_y is None else search_radius_y, x, np.int64, "search_radius_y")
    csx = _field(chip_size if chip_size_x is None else chip_size_x, x, np.int64, "chip_size_x")
    if chip_size_y is None:
        csy = _round_even(csx * chip_aspect).astype(np.int64)
    else:
        csy = _field(chip_size_y, x, np.int64, "chip_size_y")
    return PointSet(x, y, rx, ry, _field(dx_prior, x, np.float64, "dx_prior"),
                    _field(dy_prior, x, np.float64, "dy_prior"), csx, csy)


def gridpoints_xy(xs, ys, **kw):
    """Grid from explicit coordinate vectors (src/points.jl:260-265): X varies along columns."""
    X, Y = np.meshgrid(np.asarray(xs, float), np.asarray(ys, float))
    return pointset(X, Y, **kw)


def gridpoints(imagesize, spacing, chip_size=32, chip_aspect=1.0, search_radius=25,
               search_radius_x=None, search_radius_y=None, **kw):
    """Centres every `spacing` px, inset by a chip half-extent plus radius (src/points.jl:231-258)."""
    if spacing <= 0:
        raise ArgumentError(f"`spacing` must be positive, got {spacing}")
    rx0, ry0 = search_radius if isinstance(search_radius, tuple) else (search_radius, search_radius)
    rx = rx0 if search_radius_x is None else search_radius_x
    ry = ry0 if search_radius_y is None else search_radius_y
    csx = int(np.max(chip_size))
    csy = 2 * int(np.rint(csx * chip_aspect / 2))
    margin_x = -(-csx // 2) + int(np.max(rx)) + 1
    margin_y = -(-csy // 2) + int(np.max(ry)) + 1
    nrows, ncols = imagesize
    xs = np.arange(margin_x + 1, ncols - margin_x + 1, spacing)
    ys = np.arange(margin_y + 1, nrows - margin_y + 1, spacing)
    if xs.size == 0 or ys.size == 0:
        raise ArgumentError(
            f"no search center fits in a {nrows}x{ncols} image with chip {csx}x{csy} and search "
            f"radius {rx}x{ry}: a margin of {margin_y}x{margin_x} pixels is needed on each side. "
            "Use a smaller chip size or search radius.")
    return gridpoints_xy(xs, ys, chip_size=chip_size, chip_aspect=chip_aspect,
                         search_radius=search_radius, search_radius_x=search_radius_x,
                         search_radius_y=search_radius_y, **kw)


def sanitize_(pts: PointSet, min_radius: int) -> int:
    """sanitize! (src/points.jl:332-346): zero is contagious across axes; searchable points get
    at least `min_radius`. In place; returns the searchable count."""
    rx, ry = pts.radius_x, pts.radius_y
    dead = (rx <= 0) | (ry <= 0)
    rx[dead] = 0
    ry[dead] = 0
    np.maximum(rx, np.where(dead, 0, min_radius), out=rx)
    np.maximum(ry, np.where(dead, 0, min_radius), out=ry)
    return int(np.count_nonzero(~dead))


def chip_bounds(pts: PointSet, i):
    """1-based inclusive (rows, cols) of the chip at flat index i (src/points.jl:358-366)."""
    f = lambda a: np.ravel(a, order="F")[i]
    hx, hy = f(pts.chip_size_x) // 2, f(pts.chip_size_y) // 2
    cx = math.floor(f(pts.x) - f(pts.dx_prior))
    cy = math.floor(f(pts.y) - f(pts.dy_prior))
    return (cy - hy, cy + hy - 1), (cx - hx, cx + hx - 1)


def search_bounds(pts: PointSet, i):
    """1-based inclusive (rows, cols) of the search window (src/points.jl:385-395)."""
    f = lambda a: np.ravel(a, order="F")[i]
    hx, hy = f(pts.chip_size_x) // 2, f(pts.chip_size_y) // 2
    rx, ry = f(pts.radius_x), f(pts.radius_y)
    cx, cy = math.floor(f(pts.x)), math.floor(f(pts.y))
    return (cy - hy - ry, cy + hy + ry - 2), (cx - hx - rx, cx + hx + rx - 2)
