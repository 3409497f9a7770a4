"""GPU parity of the direct Float64 tracker (Float32 images, and UInt8 chips whose area is not a
power of two) against the oracle, bit for bit, on the geometry that stresses it: chip heights
that are not a multiple of 8, radius 40, chip 128, chips whose values span so many binades that
their mean must be summed in the reference's order, windows too large for its LDS (the pass then
takes the one-kernel generic tracker), and equality with that generic tracker.
"""
import os

import numpy as np
import pytest

import oracle
from _util import shifted_pair

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _check(ctx, ref, sec, ok, pts, up, path, measure=oracle.ZNCC):
    ctx.set_pair(ref, sec, ok, ok)
    g = ctx.track(pts, upsampling=up, measure=measure)
    assert ctx.last_kernel()[1] == path
    e = oracle.track(ref, sec, ok, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                     pts.chip_size_y, pts.dx_prior, pts.dy_prior, up=up, measure=measure,
                     nthreads=8)
    np.testing.assert_array_equal(g.searched, e[3])
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(_bits(a), _bits(b))
    return g


@pytest.fixture(scope="module")
def ctx():
    c = autorift.Context()
    yield c
    c.close()


@pytest.mark.parametrize("cs,csy,r", [(32, 32, 25), (32, 26, 25), (48, 48, 12), (64, 64, 40),
                                      (128, 128, 25), (16, 18, 7), (40, 24, 31)])
@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_direct_shapes(ctx, cs, csy, r, dtype):
    n = 320
    ref, sec = shifted_pair(n, (2.25, -3.5), seed=cs + r, dtype=dtype)
    if dtype == np.uint8 and (cs * csy) & (cs * csy - 1) == 0:
        pytest.skip("power-of-two chip areas of UInt8 pairs take the integer matrix-core path")
    ok = np.ones(ref.shape, bool)
    rng = np.random.default_rng(cs * r)
    m = 40
    x = rng.uniform(10, n - 10, m)
    y = rng.uniform(10, n - 10, m)
    pts = autorift.pointset(x, y, search_radius=r, chip_size_x=cs, chip_size_y=csy,
                            dx_prior=rng.normal(0, 2, m), dy_prior=rng.normal(0, 2, m))
    print(f"direct f64 chip {cs}x{csy} r{r} {np.dtype(dtype)}", flush=True)
    for up in (1, 64):
        _check(ctx, ref, sec, ok, pts, up, 3)
    _check(ctx, ref, sec, ok, pts, 16, 3, measure=oracle.NCC)


def test_direct_mixed_sizes_masks_and_degenerate(ctx):
    ref, sec = shifted_pair(256, (-1.75, 2.5), seed=11, dtype=np.float32)
    ref = ref.copy()
    sec = sec.copy()
    ok = np.ones(ref.shape, bool)
    ok[30:60, 100:150] = False
    ref[200:, :50] = 3.0          # flat: degenerate chips
    sec[200:, :50] = 3.0
    rng = np.random.default_rng(3)
    m = 200
    pts = autorift.pointset(rng.uniform(-10, 266, m), rng.uniform(-10, 266, m),
                            search_radius_x=rng.integers(0, 30, m),
                            search_radius_y=rng.integers(0, 30, m),
                            chip_size_x=rng.choice([16, 24, 32, 48], m),
                            chip_size_y=rng.choice([16, 24, 32, 48], m))
    print("direct f64 mixed", flush=True)
    for up in (1, 64):
        _check(ctx, ref, sec, ok, pts, up, 3)


def test_direct_sequential_mean(ctx):
    """Chips whose values span ~60 binades: the mean's sum is not exact in any order, so it runs
    in one lane in the reference's order; the result is still bit-exact."""
    ref, sec = shifted_pair(200, (1.5, -0.5), seed=12, dtype=np.float32)
    rng = np.random.default_rng(5)
    tiny = rng.random(ref.shape) < 0.3
    ref = np.where(tiny, ref * np.float32(1e-12), ref * np.float32(1e6)).astype(np.float32)
    sec = np.where(tiny, sec * np.float32(1e-12), sec * np.float32(1e6)).astype(np.float32)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(50.0, 151.0, 20.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=12)
    print("direct f64 sequential mean", flush=True)
    _check(ctx, ref, sec, ok, pts, 64, 3)


def test_direct_equals_generic_and_large_window_fallback(ctx):
    ref, sec = shifted_pair(300, (3.25, 1.5), seed=13, dtype=np.float32)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(60.0, 241.0, 30.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=20)
    ctx.set_pair(ref, sec, ok, ok)
    a = ctx.track(pts, upsampling=64)
    assert ctx.last_kernel()[1] == 3
    os.environ["AR_F64_DIRECT"] = "never"
    try:
        b = ctx.track(pts, upsampling=64)
        assert ctx.last_kernel()[1] == 1
    finally:
        del os.environ["AR_F64_DIRECT"]
    for u, v in zip((a.dx, a.dy, a.correlation), (b.dx, b.dy, b.correlation)):
        np.testing.assert_array_equal(_bits(u), _bits(v))
    # chip 128 at radius 64: the window (255 columns) does not fit the direct tracker's LDS
    big = autorift.gridpoints_xy(np.array([150.0]), np.array([150.0]), chip_size=128,
                                 search_radius=64)
    print("direct f64 fallback", flush=True)
    ref2, sec2 = shifted_pair(300, (3, -2), seed=14, dtype=np.float32)
    _check(ctx, ref2, sec2, ok, big, 64, 1)


def test_direct_chunks_on_two_streams_equal_generic(ctx):
    """20000 points at chip 64, radius 40 (about 135 KB of staging each): several chunks, prepared
    on the second stream while the previous chunk is tracked. Every bit equals the one-kernel
    generic tracker (itself bit-exact against the oracle above)."""
    ref, sec = shifted_pair(1200, (1.25, -2.75), seed=15, dtype=np.float32)
    ok = np.ones(ref.shape, bool)
    rng = np.random.default_rng(8)
    m = 20000
    pts = autorift.pointset(rng.uniform(80, 1120, m), rng.uniform(80, 1120, m),
                            search_radius=40, chip_size=64)
    ctx.set_pair(ref, sec, ok, ok)
    print("direct f64 chunks", flush=True)
    a = ctx.track(pts, upsampling=64)
    assert ctx.last_kernel()[1] == 3
    os.environ["AR_F64_DIRECT"] = "never"
    try:
        b = ctx.track(pts, upsampling=64)
    finally:
        del os.environ["AR_F64_DIRECT"]
    np.testing.assert_array_equal(a.searched, b.searched)
    for u, v in zip((a.dx, a.dy, a.correlation), (b.dx, b.dy, b.correlation)):
        np.testing.assert_array_equal(_bits(u), _bits(v))
    assert np.isfinite(a.dx).mean() > 0.99
