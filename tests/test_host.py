"""CPU tests of the host side: the C ABI library loads and exports what include/*.h declares,
parameter resolution, and point-set geometry (test/params.jl, test/points.jl)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import autorift
from autorift import ArgumentError, DimensionMismatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(ar_[a-z_0-9]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = autorift.lib()
    declared = _declared()
    assert {"ar_context_create", "ar_track", "ar_run_multichip", "ar_set_images"} <= declared
    for name in sorted(declared):
        assert hasattr(lib, name), name
    bound = {s[0] for s in autorift._lib.SIGNATURES}
    assert declared == bound


def test_version_and_defaults_without_device():
    assert autorift.lib().ar_version().startswith(b"autorift_hip")
    p = autorift.default_params()
    assert (p.chip_size_base, p.chip_size_max, p.upsampling, p.search_radius_x) == (32, 128, 64, 25)
    assert (p.gardner_window, p.gardner_iterations, p.fill_window) == (5, 3, 3)
    assert abs(p.min_agree_fraction - 8 / 25) < 1e-15


def test_no_device_is_a_status_not_a_crash():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(autorift.NoDeviceError):
        autorift.Context()


def test_params_defaults_and_validation():
    p = autorift.params()
    assert (p.chip_size_base, p.chip_size_min, p.chip_size_max, p.grid_spacing) == (32, 32, 128, 32)
    assert p.preprocess == 1 and p.quantize == 1 and p.upsampling == 64 and p.outliers == 1
    assert autorift.chip_sizes(p) == [32, 64, 128]
    assert autorift.chip_sizes(autorift.params(chip_size=16, chip_size_max=128)) == [16, 32, 64, 128]
    assert autorift.params(subpixel="none").upsampling == 1
    assert autorift.params(search_radius=(10, 20)).search_radius_y == 20
    for bad in (dict(chip_size=30), dict(chip_size=32, chip_size_max=96),
                dict(chip_size=32, chip_size_min=16), dict(upsampling=48),
                dict(similarity="cosine"), dict(outlier_window=4), dict(fill_window=2),
                dict(search_radius=0), dict(outliers="none", mad_scale=3.0),
                dict(min_coarse_valid_fraction=1.5)):
        with pytest.raises(ArgumentError):
            autorift.params(**bad)


def test_preprocess_methods_and_widths():
    """Highpass / Wallis as symbols with filter_width or as method instances; widths follow
    _check_filter_width (src/types.jl:587-594); windowless methods ignore filter_width
    (src/params.jl:91-103)."""
    p = autorift.params(preprocess="wallis", filter_width=21)
    assert (p.preprocess, p.filter_width, p.min_std) == (3, 21, 0.0)
    p = autorift.params(preprocess=autorift.Wallis(width=7, min_std=0.25))
    assert (p.preprocess, p.filter_width, p.min_std) == (3, 7, 0.25)
    p = autorift.params(preprocess=autorift.Highpass(width=9))
    assert (p.preprocess, p.filter_width) == (1, 9)
    assert autorift.params(preprocess="none", filter_width=4).preprocess == 0
    assert autorift.params().min_std == 0.0 and autorift.params().filter_width == 5
    for bad in (dict(preprocess="wallis", filter_width=4), dict(preprocess="wallis", filter_width=1),
                dict(preprocess="highpass", filter_width=6), dict(preprocess="sobel")):
        with pytest.raises(ArgumentError):
            autorift.params(**bad)
    for bad in (dict(width=4), dict(width=1), dict(min_std=-0.1), dict(min_std=float("nan"))):
        with pytest.raises(ArgumentError):
            autorift.Wallis(**bad)
    with pytest.raises(ArgumentError):
        autorift.Highpass(width=2)


def test_pointset_and_grid_geometry():
    pts = autorift.pointset([10.0, 50.5, 200.0], [20.0, 60.0, 30.25], search_radius=25,
                            chip_size=32)
    assert len(pts) == 3 and pts.radius_x[1] == 25 and pts.chip_size_x[1] == 32
    g = autorift.gridpoints((512, 512), 32, chip_size=32, search_radius=25)
    # margin cld(32, 2) + 25 + 1 = 42 and centres 43:32:470 (src/points.jl:244-249)
    assert g.shape == (14, 14)
    assert g.x[0, 0] == 43 and g.y[0, 0] == 43 and g.x[0, 1] == 75 and g.y[1, 0] == 75
    for i in range(len(g)):
        (r0, r1), (c0, c1) = autorift.search_bounds(g, i)
        assert r0 >= 1 and c0 >= 1 and r1 <= 512 and c1 <= 512
    assert autorift.gridpoints_xy(np.arange(100.0, 201.0, 10.0), np.arange(50.0, 101.0, 10.0),
                                  chip_size=16).shape == (6, 11)
    with pytest.raises(ArgumentError, match="no search center fits"):
        autorift.gridpoints((32, 32), 8, chip_size=64, search_radius=25)
    with pytest.raises(ArgumentError, match="must be positive"):
        autorift.gridpoints((512, 512), 0)
    with pytest.raises(ArgumentError):
        autorift.gridpoints((40, 40), 8, chip_size=32, search_radius=25)
    with pytest.raises(DimensionMismatch):
        autorift.pointset([1.0, 2.0], [1.0])
    with pytest.raises(ArgumentError):
        autorift.pointset([1.0], [1.0], chip_size=32.5)
    # chip_aspect rounds the height to even
    assert autorift.pointset([1.0], [1.0], chip_size=32, chip_aspect=1.3).chip_size_y[0] == 42
    # bounds: even chip centred half a pixel off, asymmetric window (src/points.jl:358-395)
    p1 = autorift.pointset([100.0], [50.0], search_radius=10, chip_size=16, dx_prior=3.0)
    assert autorift.chip_bounds(p1, 0) == ((42, 57), (89, 104))
    assert autorift.search_bounds(p1, 0) == ((32, 66), (82, 116))
    r, c = autorift.search_bounds(p1, 0)
    assert (r[1] - r[0] + 1) == 16 + 2 * 10 - 1


def test_sanitize():
    pts = autorift.pointset(np.zeros(4), np.zeros(4), search_radius_x=[0, 3, 10, -1],
                            search_radius_y=[5, 3, 2, 4])
    assert autorift.sanitize_(pts, 6) == 2
    assert list(pts.radius_x) == [0, 6, 10, 0] and list(pts.radius_y) == [0, 6, 6, 0]


def _reflect101_0(p, n):
    """src/peak.jl:374-382, literally."""
    if n == 1:
        return 0
    m = n - 1
    while p < 0 or p > m:
        if p < 0:
            p = -p
        if p > m:
            p = 2 * m - p
    return p


def _taps_reference(k, n):
    """src/peak.jl:312-324 with 0-based k and 0-based source indices."""
    p = k
    if p % 2 == 0:
        return (_reflect101_0(p - 2, 2 * n) // 2, p // 2, _reflect101_0(p + 2, 2 * n) // 2, 3)
    c = _reflect101_0(p - 1, 2 * n) // 2
    return (c, c, _reflect101_0(p + 1, 2 * n) // 2, 2)


def _taps_closed(k, n):
    """The closed form the device cascade uses (csrc/track_common.h: taps)."""
    c = k >> 1
    q = min(c + 1, n - 1)
    if k % 2 == 0:
        return (1 if c == 0 else c - 1, c, q, 3)
    return (c, c, q, 2)


def test_pyrup_taps_closed_form():
    for n in list(range(2, 64)) + [80, 160, 320, 640]:
        for k in range(2 * n):
            assert _taps_closed(k, n) == _taps_reference(k, n), (k, n)


def test_rotation_search_construction():
    """RotationSearch and params(rotation=...) (test/firstguess.jl:107-153)."""
    RS = autorift.RotationSearch
    assert autorift.angles(RS()) == (-3.0, 0.0, 3.0)
    assert autorift.angles(RS((-1, 0, 1))) == (-1.0, 0.0, 1.0)
    assert autorift.angles(RS(about=8.0)) == (-11.0, -8.0, -5.0)
    assert autorift.angles(RS((-1, 0, 1), about=-4.5)) == (3.5, 4.5, 5.5)
    assert autorift.angles(RS((0.0,), about=-8.0)) == (8.0,)
    for bad in ((), (0.0, 3.0, 0.0), (float("nan"), 0.0, 3.0), (-float("inf"), 0.0, 3.0)):
        with pytest.raises(autorift.ArgumentError):
            RS(bad)
    with pytest.raises(autorift.ArgumentError):
        RS(about=float("nan"))
    with pytest.raises(autorift.ArgumentError):
        RS((-3, 0, 3), about=float("inf"))
    with pytest.raises(TypeError):
        RS(about=[1.0])
    assert autorift.params().rotation_count == 0
    assert autorift.params(rotation=None).rotation_count == 0
    assert autorift.params(rotation=False).rotation_count == 0
    p = autorift.params(rotation=RS((-6, -3, 0, 3, 6), about=1.0))
    assert p.rotation_count == 5
    assert list(p.rotation_angles)[:5] == [-7.0, -4.0, -1.0, 2.0, 5.0]
    with pytest.raises(autorift.ArgumentError):
        autorift.params(rotation=RS(tuple(range(9))))
    with pytest.raises(autorift.ArgumentError):
        autorift.params(rotation="yes")


def test_similarity_tuple_maps_onto_levels():
    """test/complex.jl:239-266: a scalar applies everywhere; a tuple assigns in order with the
    last entry repeating; more measures than levels and an empty tuple are errors."""
    ar = autorift
    p = ar.params(chip_size=32, chip_size_max=128)
    assert len(ar.chip_sizes(p)) == 3
    assert ar.chip_measures(p) == ("zncc", "zncc", "zncc")
    p2 = ar.params(chip_size=32, chip_size_max=128, similarity=("coherence", "zncc"))
    assert ar.chip_measures(p2) == ("coherence", "zncc", "zncc")
    assert ar.measure_at(p2, 1) == "coherence" and ar.measure_at(p2, 3) == "zncc"
    p3 = ar.params(chip_size=32, chip_size_max=128, similarity=("coherence", "ncc", "zncc"))
    assert ar.chip_measures(p3) == ("coherence", "ncc", "zncc")
    with pytest.raises(ar.ArgumentError):
        ar.params(chip_size=32, chip_size_max=32, similarity=("coherence", "zncc"))
    with pytest.raises(ar.ArgumentError):
        ar.params(similarity=())
    with pytest.raises(ar.ArgumentError):
        ar.params(deramp_axis="z")
    d = ar.params(preprocess="deramp", deramp_axis="y", quantize="none")
    assert (d.preprocess, d.deramp_axis) == (2, 2)
