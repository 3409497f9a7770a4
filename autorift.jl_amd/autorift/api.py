"""Public API with AutoRIFT.jl's shape: params, ImagePair, init / reinit_ / autorift_,
autorift, track (src/api.jl, src/params.jl, src/track.jl).

Everything numeric runs on the device through the C ABI; this module resolves keywords,
builds point sets and keeps the cache lifecycle, as the Julia host does.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, Params
from .core import Context, DisplacementField, MultichipResult
from .points import PointSet, gridpoints

_SIMILARITY = {"zncc": _lib.AR_ZNCC, "ncc": _lib.AR_NCC, "coherence": _lib.AR_COHERENCE}
_PREPROCESS = {"highpass": 1, "none": 0, "deramp": 2, "wallis": 3}
_DERAMP_AXIS = {"both": 0, "x": 1, "y": 2}
_QUANTIZE = {"uint8": 1, "none": 0}
_SUBPIXEL = ("pyramid", "none")
_OUTLIERS = ("gardner", "none")
# arithmetic of the correlation contraction for Float32 prepared images (no reference
# counterpart: the reference always correlates in Float32 / Float64)
_PRECISION = {"exact": 0, "bf16": 1, "fp16": 2, "fft": 3}


class RotationSearch:
    """Chip rotations tried at every point, the strongest peak kept (src/types.jl:426-552):
    `angles` is the search window in degrees, `about` the scene rotation it is centred on."""

    def __init__(self, angles=(-3.0, 0.0, 3.0), about=0.0):
        if isinstance(about, (bool, np.bool_)) or not isinstance(about, (int, float, np.number)):
            raise TypeError(f"`about` must be a real number, got {type(about).__name__}")
        t = tuple(float(a) for a in angles)
        if not t:
            raise ArgumentError("`RotationSearch` needs at least one angle; use `rotation=None` "
                                "to disable.")
        if len(set(t)) != len(t):
            raise ArgumentError(f"`RotationSearch` angles must be distinct, got {t} -- a repeated "
                                "angle costs a full correlation and can never win.")
        a = float(about)
        if not (all(np.isfinite(t)) and np.isfinite(a)):
            raise ArgumentError(f"`RotationSearch` needs finite angles and a finite `about`, got "
                                f"angles {t} about {a}.")
        self.angles = t
        self.about = a

    def __repr__(self):
        return f"RotationSearch({self.angles}, about={self.about})"


def _check_filter_width(width, who):
    """src/types.jl:587-594: a window filter's width is odd and at least 3."""
    w = int(width)
    if w < 3:
        raise ArgumentError(f"{who} `width` must be >= 3, got {w}")
    if w % 2 == 0:
        raise ArgumentError(f"{who} `width` must be odd so the filter has a well-defined centre, "
                            f"got {w}")
    return w


class Highpass:
    """Highpass(; width = 5) (src/types.jl:286-297): subtract the local box mean."""

    def __init__(self, width=5):
        self.width = _check_filter_width(width, "Highpass")

    def __repr__(self):
        return f"Highpass(width={self.width})"


class Wallis:
    """Wallis(; width = 5, min_std = 0.0) (src/types.jl:299-322): subtract the local mean and
    divide by the local standard deviation, floored at `min_std` (0 disables the floor)."""

    def __init__(self, width=5, min_std=0.0):
        ms = float(min_std)
        if not ms >= 0:
            raise ArgumentError(f"`min_std` must be >= 0, got {min_std}")
        self.width = _check_filter_width(width, "Wallis")
        self.min_std = ms

    def __repr__(self):
        return f"Wallis(width={self.width}, min_std={self.min_std})"


def angles(m: RotationSearch):
    """The chip rotations a pass tries, `about` subtracted (src/types.jl:577)."""
    return tuple(a - m.about for a in m.angles)


def _resolve(table, x, kw):
    if isinstance(x, str) and x.lstrip(":") in table:
        return x.lstrip(":")
    valid = ", ".join(f":{k}" for k in sorted(table))
    raise ArgumentError(f"`{kw} = {x!r}` is not recognised. Valid options are {valid}.")


def _check_chip_size(name, cs):
    cs = int(cs)
    if cs <= 0:
        raise ArgumentError(f"`{name}` must be positive, got {cs}")
    if cs % 4:
        raise ArgumentError(f"`{name}` must be a multiple of 4, got {cs}. Chip sizes are "
                            "halved at each chip-size level and need an even half-extent in "
                            "both directions.")
    return cs


def _pair(x):
    if isinstance(x, (tuple, list)):
        if len(x) != 2:
            raise ArgumentError(f"expected 1 or 2 values, got {len(x)}")
        return int(x[0]), int(x[1])
    return int(x), int(x)


def params(similarity="zncc", preprocess="highpass", quantize="uint8", subpixel="pyramid",
           filter_width=None, upsampling=None, chip_size=32, chip_size_min=None,
           chip_size_max=None, chip_aspect=1.0, grid_spacing=None, search_radius=25,
           search_radius_x=None, search_radius_y=None, min_search_radius=6, dx_prior=0.0,
           dy_prior=0.0, coarse_stride=4, coarse_buffer=8, min_coarse_valid_fraction=0.01,
           outliers="gardner", rotation=None, outlier_window=None, outlier_iterations=None,
           min_agree_fraction=None, agree_tolerance=None, mad_scale=None, fill_window=3,
           threaded=False, progress=False, rng_seed=0, precision="exact",
           deramp_axis="both") -> Params:
    """Resolve and validate keywords into the flattened parameter set (src/params.jl:306-392).

    `precision` (:exact, :bf16, :fp16) is this library's addition: with `quantize = :none` the
    correlation contraction runs on the matrix cores with 16-bit operands and FP32 accumulation;
    UInt8 prepared images always take the exact integer path.

    `similarity` is a measure or a tuple of measures, one per chip-size level with the last
    repeating (src/types.jl:770-870), e.g. ("coherence", "zncc"). `preprocess = "deramp"` with
    `deramp_axis` ("both", "x", "y") is Deramp (src/types.jl:383-423), for complex pairs.
    """
    p = _lib.default_params()
    p.precision = _PRECISION[_resolve(_PRECISION, precision, "precision")]
    sims = tuple(similarity) if isinstance(similarity, (tuple, list)) else (similarity,)
    if not sims:
        raise ArgumentError("`similarity` must name at least one measure")
    if len(sims) > 8:
        raise ArgumentError(f"at most 8 per-level measures, got {len(sims)}")
    meas = [_SIMILARITY[_resolve(_SIMILARITY, s, "similarity")] for s in sims]
    p.measure = meas[0]
    if len(meas) > 1:
        p.level_measure_count = len(meas)
        for k, v in enumerate(meas):
            p.level_measures[k] = v
    p.deramp_axis = _DERAMP_AXIS[_resolve(_DERAMP_AXIS, deramp_axis, "deramp_axis")]
    if isinstance(preprocess, (Highpass, Wallis)):
        # a method instance carries its own width (src/params.jl:90)
        p.preprocess = _PREPROCESS["wallis" if isinstance(preprocess, Wallis) else "highpass"]
        p.filter_width = preprocess.width
        if isinstance(preprocess, Wallis):
            p.min_std = preprocess.min_std
    else:
        pre = _resolve(_PREPROCESS, preprocess, "preprocess")
        p.preprocess = _PREPROCESS[pre]
        # windowless methods ignore filter_width (src/params.jl:91-103)
        if filter_width is not None and pre in ("highpass", "wallis"):
            p.filter_width = _check_filter_width(filter_width, pre.capitalize())
    p.quantize = _QUANTIZE[_resolve(_QUANTIZE, quantize, "quantize")]
    sub = _resolve({k: 0 for k in _SUBPIXEL}, subpixel, "subpixel")
    if sub == "none":
        p.upsampling = 1
    else:
        up = 64 if upsampling is None else int(upsampling)
        if up <= 1:
            raise ArgumentError(f"`upsampling` must be > 1, got {up}")
        if up & (up - 1):
            raise ArgumentError(f"`upsampling` must be a power of 2, got {up}")
        p.upsampling = up
    out = _resolve({k: 0 for k in _OUTLIERS}, outliers, "outliers")
    gkw = dict(window=outlier_window, iterations=outlier_iterations,
               min_agree_fraction=min_agree_fraction, agree_tolerance=agree_tolerance,
               mad_scale=mad_scale)
    if out == "none":
        given = [k for k, v in gkw.items() if v is not None]
        if given:
            raise ArgumentError(f"`outliers` was given `:none`, which takes no parameters, so "
                                f"{', '.join(given)} cannot also apply.")
        p.outliers = 0
    else:
        p.outliers = 1
        w = 5 if outlier_window is None else int(outlier_window)
        if w % 2 == 0:
            raise ArgumentError(f"outlier filter `window` must be odd so the neighbourhood has "
                                f"a centre, got {w}")
        if w < 3:
            raise ArgumentError(f"outlier filter `window` must be >= 3, got {w}")
        it = 3 if outlier_iterations is None else int(outlier_iterations)
        if it < 1:
            raise ArgumentError(f"outlier filter `iterations` must be >= 1, got {it}")
        maf = 8 / 25 if min_agree_fraction is None else float(min_agree_fraction)
        if not 0 <= maf <= 1:
            raise ArgumentError(f"`min_agree_fraction` must be in [0, 1], got {maf}")
        at = 0.2 if agree_tolerance is None else float(agree_tolerance)
        if not 0 < at <= 1:
            raise ArgumentError(f"`agree_tolerance` must be in [0, 1] and non-zero, got {at}")
        ms = 4.0 if mad_scale is None else float(mad_scale)
        if not ms > 0:
            raise ArgumentError(f"`mad_scale` must be positive, got {ms}")
        p.gardner_window, p.gardner_iterations = w, it
        p.min_agree_fraction, p.agree_tolerance, p.mad_scale = maf, at, ms
    if rotation is not None and rotation is not False:
        if not isinstance(rotation, RotationSearch):
            raise ArgumentError(f"`rotation` must be a RotationSearch, None or False, got "
                                f"{rotation!r}")
        ang = angles(rotation)
        if len(ang) > 8:
            raise ArgumentError(f"at most 8 rotation angles on the device path, got {len(ang)}")
        p.rotation_count = len(ang)
        for k, v in enumerate(ang):
            p.rotation_angles[k] = v
    base = _check_chip_size("chip_size", chip_size)
    cmin = base if chip_size_min is None else _check_chip_size("chip_size_min", chip_size_min)
    cmax = 4 * base if chip_size_max is None else _check_chip_size("chip_size_max",
                                                                   chip_size_max)
    if cmin > cmax:
        raise ArgumentError(f"`chip_size_min` ({cmin}) must be <= `chip_size_max` ({cmax})")
    if base > cmin:
        raise ArgumentError(f"`chip_size` ({base}) must be <= `chip_size_min` ({cmin}). "
                            "`chip_size` is the finest chip-size level; coarser levels are its "
                            "powers-of-two multiples.")
    if cmax % base:
        raise ArgumentError(f"`chip_size_max` ({cmax}) must be a power-of-two multiple of "
                            f"`chip_size` ({base}), because chip-size levels double at each step.")
    ratio = cmax // base
    if ratio & (ratio - 1):
        raise ArgumentError(f"`chip_size_max` ({cmax}) must be a power-of-two multiple of "
                            f"`chip_size` ({base}), got a ratio of {ratio}.")
    p.chip_size_base, p.chip_size_min, p.chip_size_max = base, cmin, cmax
    nlev = len(chip_sizes(p))
    if len(meas) > nlev:
        raise ArgumentError(f"{len(meas)} similarity measures for {nlev} chip-size level(s): the "
                            "extra measures would silently do nothing.")
    if not chip_aspect > 0:
        raise ArgumentError(f"`chip_aspect` must be positive, got {chip_aspect}")
    p.chip_aspect = float(chip_aspect)
    if grid_spacing is None:
        p.grid_spacing = base
    else:
        if not grid_spacing > 0:
            raise ArgumentError(f"`grid_spacing` must be positive, got {grid_spacing}")
        p.grid_spacing = int(grid_spacing)
    rx0, ry0 = _pair(search_radius)
    rx = rx0 if search_radius_x is None else int(search_radius_x)
    ry = ry0 if search_radius_y is None else int(search_radius_y)
    if rx < 0:
        raise ArgumentError(f"`search_radius_x` must be >= 0, got {rx}")
    if ry < 0:
        raise ArgumentError(f"`search_radius_y` must be >= 0, got {ry}")
    if rx == 0 and ry == 0:
        raise ArgumentError("`search_radius` is zero in both axes, so no pixel can be searched.")
    p.search_radius_x, p.search_radius_y = rx, ry
    p.min_search_radius = int(min_search_radius)
    if not coarse_stride > 0:
        raise ArgumentError(f"`coarse_stride` must be positive, got {coarse_stride}")
    p.coarse_stride = int(coarse_stride)
    p.coarse_buffer = int(coarse_buffer)
    if not 0 <= min_coarse_valid_fraction <= 1:
        raise ArgumentError(f"`min_coarse_valid_fraction` must be in [0, 1], got "
                            f"{min_coarse_valid_fraction}")
    p.min_coarse_valid_fraction = float(min_coarse_valid_fraction)
    p.dx_prior, p.dy_prior = float(dx_prior), float(dy_prior)
    fw = int(fill_window)
    if fw < 1:
        raise ArgumentError(f"`fill_window` must be >= 1, got {fw}")
    if fw % 2 == 0:
        raise ArgumentError(f"`fill_window` must be odd so the window has a centre pixel, got {fw}")
    p.fill_window = fw
    return p


def chip_sizes(p: Params):
    """Levels base * 2^k within [min, max], ascending (src/types.jl:842-850)."""
    out, cs = [], p.chip_size_base
    while cs <= p.chip_size_max:
        if cs >= p.chip_size_min:
            out.append(cs)
        cs *= 2
    return out


def _build_grid(imagesize, p: Params) -> PointSet:
    """src/api.jl:339-347: margin from the largest chip level."""
    return gridpoints(imagesize, p.grid_spacing, chip_size=p.chip_size_max,
                      chip_aspect=p.chip_aspect, search_radius_x=p.search_radius_x,
                      search_radius_y=p.search_radius_y, dx_prior=p.dx_prior,
                      dy_prior=p.dy_prior)


@dataclass
class ImagePair:
    """The two images and their validity masks (src/preprocess.jl:41-79)."""
    reference: np.ndarray
    secondary: np.ndarray
    reference_valid: np.ndarray | None = None
    secondary_valid: np.ndarray | None = None

    def __post_init__(self):
        if np.shape(self.reference) != np.shape(self.secondary):
            raise DimensionMismatch(
                f"reference is {np.shape(self.reference)} but secondary is "
                f"{np.shape(self.secondary)}; the two images must be co-registered to a common "
                "grid before correlation")
        for name in ("reference_valid", "secondary_valid"):
            m = getattr(self, name)
            if m is not None and np.shape(m) != np.shape(self.reference):
                raise DimensionMismatch(f"{name} is {np.shape(m)} but the images are "
                                        f"{np.shape(self.reference)}")


class Cache:
    """Preallocated state for correlating many pairs with one configuration (src/api.jl:30-44):
    the parameters, the grid and the device context holding the prepared pair."""

    def __init__(self, p: Params, ctx: Context, grid: PointSet, raw: ImagePair):
        self.params = p
        self.ctx = ctx
        self.grid = grid
        self.raw = raw
        self.result = None
        self.isfresh = True

    def close(self):
        self.ctx.close()


def init(reference, secondary, reference_valid=None, secondary_valid=None, device=0,
         **kwargs) -> Cache:
    """Prepare to correlate (src/api.jl:66-76): keywords as params()."""
    p = params(**kwargs)
    raw = ImagePair(reference, secondary, reference_valid, secondary_valid)
    ctx = Context(p, device)
    ctx.set_images(reference, secondary, reference_valid, secondary_valid)
    grid = _build_grid(np.shape(reference), p)
    return Cache(p, ctx, grid, raw)


def reinit_(cache: Cache, reference=None, secondary=None, reference_valid=None,
            secondary_valid=None) -> Cache:
    """Point the cache at a new pair (src/api.jl:104-134). Only the images that changed are
    prepared again; an image identical (by object) to either current one is reused."""
    if reference is None and secondary is None and reference_valid is None and \
            secondary_valid is None:
        return cache
    old = cache.raw
    ref = old.reference if reference is None else reference
    sec = old.secondary if secondary is None else secondary
    if np.shape(ref) != np.shape(old.reference) or np.shape(sec) != np.shape(old.reference):
        raise DimensionMismatch(
            f"new images are {np.shape(ref)} and {np.shape(sec)} but the cache was built for "
            f"{np.shape(old.reference)}; the grid and output arrays are sized to that. Build a "
            "new cache with `init` for a different image size.")

    def maskfor(img):   # a mask follows its image (src/api.jl:321-323)
        if img is old.reference:
            return old.reference_valid
        if img is old.secondary:
            return old.secondary_valid
        return None

    rv = maskfor(ref) if reference_valid is None else reference_valid
    sv = maskfor(sec) if secondary_valid is None else secondary_valid
    new = ImagePair(ref, sec, rv, sv)
    same_ref = ref is old.reference and rv is old.reference_valid
    same_sec = sec is old.secondary and sv is old.secondary_valid
    if ref is old.secondary and sec is not old.secondary:
        # a walk along a series: the new reference is the old secondary; prepare both slots
        cache.ctx.set_images(ref, sec, rv, sv)
    else:
        cache.ctx.set_images(None if same_ref else ref, None if same_sec else sec,
                             None if same_ref else rv, None if same_sec else sv)
    cache.raw = new
    cache.isfresh = True
    cache.result = None
    return cache


def autorift_(cache: Cache) -> MultichipResult:
    """Correlate the cache's current pair (src/api.jl:165-170); idempotent until reinit_."""
    if not cache.isfresh and cache.result is not None:
        return cache.result
    cache.result = cache.ctx.run_multichip(cache.grid)
    cache.isfresh = False
    return cache.result


def autorift(reference, secondary, grid: PointSet | None = None, reference_valid=None,
             secondary_valid=None, device=0, **kwargs):
    """Displacement of `secondary` relative to `reference` on a grid (src/api.jl:210-274).

    With a gridded PointSet the multi-chip-size search runs over it; with a scattered one a
    single tracking pass runs (src/api.jl:273-274). dx, dy are the offset from secondary back
    to reference (the negative of feature motion), dy increasing downward."""
    cache = init(reference, secondary, reference_valid, secondary_valid, device, **kwargs)
    try:
        if grid is None:
            return autorift_(cache)
        if len(grid.shape) == 2:
            return cache.ctx.run_multichip(grid)
        return cache.ctx.track(grid, upsampling=cache.params.upsampling,
                               measure=cache.params.measure)
    finally:
        cache.close()


def autorift_with_grid(reference, secondary, **kwargs):
    """autorift plus the grid it ran on (src/api.jl:288-295)."""
    p = params(**{k: v for k, v in kwargs.items()
                  if k not in ("reference_valid", "secondary_valid", "device")})
    grid = _build_grid(np.shape(reference), p)
    return autorift(reference, secondary, grid, **kwargs), grid


def track(pair, pts: PointSet, p: Params | None = None, subpixel=None, measure=None,
          device=0) -> DisplacementField:
    """track (src/track.jl:103-160) on an ImagePair used as is, or on a Cache's prepared pair."""
    p = p if p is not None else params()
    up = p.upsampling if subpixel is None else (1 if subpixel in ("none", 1) else int(subpixel))
    m = p.measure if measure is None else _SIMILARITY[_resolve(_SIMILARITY, measure, "measure")]
    if isinstance(pair, Cache):
        return pair.ctx.track(pts, upsampling=up, measure=m)
    with Context(p, device) as ctx:
        ctx.set_pair(pair.reference, pair.secondary, pair.reference_valid, pair.secondary_valid)
        return ctx.track(pts, upsampling=up, measure=m)


def measure_at(p: Params, k: int) -> str:
    """The measure of chip-size level k (1-based), the tuple's last entry repeating."""
    names = {v: n for n, v in _SIMILARITY.items()}
    if p.level_measure_count > 0:
        return names[p.level_measures[min(k, p.level_measure_count) - 1]]
    return names[p.measure]


def chip_measures(p: Params):
    """One measure per chip-size level (src/types.jl:852-870)."""
    return tuple(measure_at(p, k) for k in range(1, len(chip_sizes(p)) + 1))
