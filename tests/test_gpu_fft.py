"""GPU tests of the FFT tracker (precision = :fft): the reference's transform numerator
(src/correlate.jl:816-848) in single precision, its dispatch threshold (:785-797), and the exact
direct fallback for transforms the kernel does not take.

The transform's rounding is not FFTW's, so the bar is the one the reference sets for its own two
numerator modes (test/correlate.jl: peaks equal, values within 1e-4): integer peaks equal to the
oracle's FFT mode and to the exact surface except at near-ties of the exact surface, surface
values within 1e-5, subpixel offsets within 1e-3 px.
"""
import numpy as np
import pytest

import oracle
from bench import cfg1_pair
from test_gpu_configs import _cfg1_grid, _top2_gap

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")

PATH_EXACT_DIRECT, PATH_FFT = 3, 4


def test_fft_cfg1_against_oracle_fft_and_exact():
    a, b = cfg1_pair(1)
    grid = _cfg1_grid()
    res = {}
    for prec in ("fft", "exact"):
        with autorift.Context(autorift.params(quantize="none", precision=prec)) as ctx:
            ctx.set_images(a, b)
            ref, sec, rv, sv = ctx.get_pair()
            res[prec] = (ctx.track(grid, upsampling=1), ctx.track(grid, upsampling=64),
                         ctx.last_kernel()[1])
    g1, g64, path = res["fft"]
    e1, e64, _ = res["exact"]
    assert path == PATH_FFT
    ok = rv & sv
    args = (ref, sec, ok, grid.x, grid.y, 25, 25, 32, 32)
    f1 = oracle.track(*args, up=1, num_mode=oracle.NUM_FFT, nthreads=16)
    f64 = oracle.track(*args, up=64, num_mode=oracle.NUM_FFT, nthreads=16)
    meas = np.isfinite(g1.dx)
    np.testing.assert_array_equal(meas, np.isfinite(f1[0]))
    np.testing.assert_array_equal(meas, np.isfinite(e1.dx))
    near = _top2_gap(ref, sec, ok, grid) < 1e-5
    for name, (ox, oy) in (("oracle fft", (f1[0], f1[1])), ("exact", (e1.dx, e1.dy))):
        differ = meas & ((g1.dx != ox) | (g1.dy != oy))
        print(f"fft vs {name}: {meas.sum()} points, integer peaks differing {int(differ.sum())}, "
              f"near-ties {int((near & meas).sum())}", flush=True)
        assert not (differ & ~near).any()
    np.testing.assert_allclose(g1.correlation[meas], e1.correlation[meas], rtol=0, atol=1e-5)
    np.testing.assert_allclose(g1.correlation[meas], f1[2][meas], rtol=0, atol=1e-5)
    for name, (ox, oy) in (("oracle fft", (f64[0], f64[1])), ("exact", (e64.dx, e64.dy))):
        err = np.hypot(g64.dx - ox, g64.dy - oy)[meas]
        print(f"fft vs {name}: subpixel |d| > 1e-3 px at {int((err > 1e-3).sum())} of {err.size}, "
              f"max {err.max():.3g} px", flush=True)
        assert (err <= 1e-3).mean() >= 0.99
        assert err.max() <= 1 / 64 * np.sqrt(2) + 1e-6
    # recovery of the known shift (outputs are secondary -> reference offsets)
    assert abs(np.median(g64.dx[meas]) + 3.30) < 0.05 and abs(np.median(g64.dy[meas]) - 1.70) < 0.05


@pytest.mark.parametrize("cs,csy,r,ry", [(16, 16, 8, 8), (32, 32, 16, 25), (64, 64, 25, 25),
                                         (48, 32, 20, 12), (32, 32, 40, 40), (24, 40, 30, 10)])
def test_fft_shapes_against_exact(cs, csy, r, ry):
    """Transform sizes from every radix mix the default next_fft_size picks; non-square chips
    and radii. Against the exact direct surface of the same pair."""
    rng = np.random.default_rng(cs + 7 * r)
    a, b = cfg1_pair(20 + cs, n=400, shift=(rng.uniform(-4, 4), rng.uniform(-4, 4)))
    xs = np.arange(80.0, 321.0, 12.0)
    X, Y = np.meshgrid(xs, xs)
    pts = autorift.pointset(X.ravel(), Y.ravel(), search_radius_x=r, search_radius_y=ry,
                            chip_size_x=cs, chip_size_y=csy)
    out = {}
    for prec in ("fft", "exact"):
        with autorift.Context(autorift.params(quantize="none", precision=prec)) as ctx:
            ctx.set_images(a, b)
            out[prec] = (ctx.track(pts, upsampling=1), ctx.last_kernel()[1])
    (g, path), (e, _) = out["fft"], out["exact"]
    work = (2 * r) * (2 * ry) * cs * csy
    print(f"fft shape chip {cs}x{csy} r {r}x{ry}: path {path} (direct below 20000: {work})",
          flush=True)
    assert path == (PATH_FFT if work > 20000 else PATH_EXACT_DIRECT)
    meas = np.isfinite(e.dx)
    np.testing.assert_array_equal(np.isfinite(g.dx), meas)
    same = (g.dx[meas] == e.dx[meas]) & (g.dy[meas] == e.dy[meas])
    print(f"  integer peaks equal {same.mean():.4f} over {meas.sum()}", flush=True)
    assert same.mean() >= 0.99
    np.testing.assert_allclose(g.correlation[meas][same], e.correlation[meas][same], rtol=0,
                               atol=1e-5)


def test_fft_dispatch_below_threshold_and_large_transforms():
    """Below DIRECT_THRESHOLD (20 000 multiply-adds) the reference correlates directly, and so
    does this path: results bit-equal to precision = :exact. Transforms above 128 per side (chip
    128, r 25: window 177) take the exact direct tracker too."""
    a, b = cfg1_pair(3, n=600)
    xs = np.arange(150.0, 451.0, 30.0)
    X, Y = np.meshgrid(xs, xs)
    for cs, r in ((8, 4), (128, 25)):
        pts = autorift.pointset(X.ravel(), Y.ravel(), search_radius=r, chip_size=cs)
        out = {}
        for prec in ("fft", "exact"):
            with autorift.Context(autorift.params(quantize="none", precision=prec)) as ctx:
                ctx.set_images(a, b)
                out[prec] = (ctx.track(pts, upsampling=64), ctx.last_kernel()[1])
        (g, path), (e, _) = out["fft"], out["exact"]
        print(f"fft dispatch chip {cs} r {r}: path {path}", flush=True)
        assert path == PATH_EXACT_DIRECT
        for x, y in ((g.dx, e.dx), (g.dy, e.dy), (g.correlation, e.correlation)):
            np.testing.assert_array_equal(np.asarray(x, np.float32).view(np.uint32),
                                          np.asarray(y, np.float32).view(np.uint32))


def test_fft_multichip_default_levels():
    """The default multi-chip search (levels 32, 64 by transform, 128 exact) on the Float32
    cfg1 pair: about the same measured points as the exact search, 99 % of offsets within
    1e-3 px of it, the known shift recovered."""
    a, b = cfg1_pair(1)
    grid = _cfg1_grid()
    res = {}
    for prec in ("fft", "exact"):
        with autorift.Context(autorift.params(quantize="none", precision=prec)) as ctx:
            ctx.set_images(a, b)
            res[prec] = ctx.run_multichip(grid)
    g, e = res["fft"], res["exact"]
    ok = np.isfinite(g.dx) & np.isfinite(e.dx)
    print(f"fft multichip: measured {np.isfinite(g.dx).sum()} vs {np.isfinite(e.dx).sum()}, "
          f"identical {(g.dx[ok] == e.dx[ok]).mean():.4f}", flush=True)
    assert abs(int(np.isfinite(g.dx).sum()) - int(np.isfinite(e.dx).sum())) <= 2
    err = np.hypot(g.dx[ok] - e.dx[ok], g.dy[ok] - e.dy[ok])
    assert (err <= 1e-3).mean() >= 0.99
    assert abs(np.nanmedian(g.dx) + 3.30) < 0.05 and abs(np.nanmedian(g.dy) - 1.70) < 0.05
