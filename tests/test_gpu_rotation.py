"""GPU parity of RotationSearch (src/track.jl:316-409) against the oracle: the generic Float64
path with the chip rotated per angle, bit for bit (sines and cosines come from the host's C
library in both)."""
import numpy as np
import pytest

import oracle
from _util import synthetic_texture

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


def _scene(n, motion, seed, dtype):
    big = synthetic_texture((n, n), seed=seed).astype(np.float32) * 255
    sec = oracle.rotate_chip(big, motion)            # the scene turned about its centre
    if dtype == np.uint8:
        return (np.clip(np.rint(big), 0, 255).astype(np.uint8),
                np.clip(np.rint(sec), 0, 255).astype(np.uint8))
    return big, sec


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("cs,r,up", [(32, 10, 1), (64, 8, 64), (48, 6, 16)])
def test_rotation_search_track(dtype, cs, r, up):
    ref, sec = _scene(320, 4.0, 11, dtype)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(100.0, 221.0, 30.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=cs, search_radius=r)
    rot = autorift.RotationSearch((-6.0, -3.0, 0.0, 3.0), about=0.5)
    p = autorift.params(preprocess="none", quantize="none", rotation=rot)
    print(f"rotation c{cs} r{r} up{up} {np.dtype(dtype).name}", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        got = ctx.track(pts, upsampling=up)
        assert ctx.last_kernel()[1] == 1
    exp = oracle.track(ref, sec, ok, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                       pts.chip_size_y, up=up, rotation=autorift.angles(rot))
    for g, e in zip((got.dx, got.dy, got.correlation), exp[:3]):
        np.testing.assert_array_equal(np.asarray(g, np.float32).view(np.uint32),
                                      np.asarray(e, np.float32).view(np.uint32))
    np.testing.assert_array_equal(got.searched, exp[3])


def test_rotation_search_recovers_correlation():
    """On a scene turned by 6 degrees, searching (-6, -3, 0) raises the median peak correlation
    over no search (the effect the reference's docstring table measures)."""
    ref, sec = _scene(400, 6.0, 12, np.float32)
    xs = np.arange(120.0, 281.0, 20.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=64, search_radius=8)
    off = autorift.track(autorift.ImagePair(ref, sec), pts,
                         autorift.params(preprocess="none", quantize="none"), subpixel=1)
    on = autorift.track(autorift.ImagePair(ref, sec), pts,
                        autorift.params(preprocess="none", quantize="none",
                                        rotation=autorift.RotationSearch((-6.0, -3.0, 0.0))),
                        subpixel=1)
    print(f"median peak {np.nanmedian(off.correlation):.3f} -> {np.nanmedian(on.correlation):.3f}")
    assert np.nanmedian(on.correlation) > np.nanmedian(off.correlation)


def test_rotation_search_multichip():
    ref, sec = _scene(300, 3.0, 13, np.uint8)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(61.0, 241.0, 12.0)
    X, Y = np.meshgrid(xs, xs)
    grid = autorift.pointset(X, Y, search_radius=8, chip_size=16)
    rot = autorift.RotationSearch((-3.0, 0.0, 3.0))
    p = autorift.params(preprocess="none", quantize="none", chip_size=16, chip_size_max=32,
                        rotation=rot)
    print("rotation multichip", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        g = ctx.run_multichip(grid)
    e = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior,
                         p=oracle.params(chip_size=16, chip_size_max=32,
                                         rotation=autorift.angles(rot)), nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])


def test_batch_refuses_rotation():
    with pytest.raises(autorift.AutoRIFTError):
        autorift.Batch(autorift.params(rotation=autorift.RotationSearch()))
