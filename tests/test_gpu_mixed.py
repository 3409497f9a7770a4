"""GPU tests of the BF16 / FP16 matrix-core tracker for Float32 prepared images.

The mixed-precision path is not bit-exact by design: the contraction runs with 16-bit operands
and FP32 accumulation. It is held to the agreement the design states, against the exact Float64
path (itself bit-exact against the oracle in test_gpu_track.py) and against the oracle directly:
  * integer peaks agree on >= 99.9 % of the points;
  * subpixel offsets agree within 0.05 px RMS; wherever the integer peaks agree they agree to
    one step of the 64x grid at most, because the 5x5 patch the cascade starts from is
    recomputed in Float64 (in a fixed parallel order, so a patch value can differ from the exact
    path's in its last Float32 bit).
"""
import numpy as np
import pytest

import oracle
from _util import shifted_pair, synthetic_texture
from bench import cfg1_pair

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")

CFG1_X = np.arange(40.0, 473.0, 16.0)   # centres 40:16:472, 784 points


def _run(prec, a, b, pts, up, measure=0, **kw):
    p = autorift.params(quantize="none", precision=prec, **kw)
    with autorift.Context(p) as ctx:
        ctx.set_images(a, b)
        r = ctx.track(pts, upsampling=up, measure=measure)
        path = ctx.last_kernel()[1]
    return r, path


def _agreement(r, e):
    both = np.isfinite(r.dx) & np.isfinite(e.dx)
    assert np.array_equal(np.isfinite(r.dx), np.isfinite(e.dx))
    assert np.array_equal(r.searched, e.searched)
    same = (r.dx[both] == e.dx[both]) & (r.dy[both] == e.dy[both])
    return same.mean() if both.any() else 1.0, both


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_cfg1_pair_agreement(prec):
    """The 512^2 Fourier-shifted pair (shift (3.30, -1.70), N(0, 2^2) noise), chip 32, r 25."""
    a, b = cfg1_pair(1)
    pts = autorift.gridpoints_xy(CFG1_X, CFG1_X, chip_size=32, search_radius=25)
    print(f"cfg1 {prec} integer peaks", flush=True)
    r1, path = _run(prec, a, b, pts, 1)
    assert path == 2
    e1, epath = _run("exact", a, b, pts, 1)
    assert epath == 3   # the direct Float64 tracker
    rate, both = _agreement(r1, e1)
    print(f"{prec}: integer-peak agreement {rate:.5f} over {both.sum()} points", flush=True)
    assert rate >= 0.999
    print(f"cfg1 {prec} subpixel", flush=True)
    r64, _ = _run(prec, a, b, pts, 64)
    e64, _ = _run("exact", a, b, pts, 64)
    ok = np.isfinite(e64.dx)
    err = np.hypot(r64.dx[ok] - e64.dx[ok], r64.dy[ok] - e64.dy[ok])
    rms = float(np.sqrt(np.mean(err ** 2)))
    print(f"{prec}: subpixel RMS vs Float64 {rms:.5f} px", flush=True)
    assert rms <= 0.05
    # where the integer peaks agree, the subpixel result is the exact path's to one 1/64 step
    same_peak = (r1.dx == e1.dx) & (r1.dy == e1.dy) & ok
    d = np.maximum(np.abs(r64.dx - e64.dx), np.abs(r64.dy - e64.dy))[same_peak]
    assert d.max() <= 1 / 64 + 1e-6
    ident = float(np.mean(d == 0))
    print(f"{prec}: subpixel identical to the Float64 path on {ident:.5f} of the points", flush=True)
    assert ident >= 0.99
    np.testing.assert_allclose(r64.correlation[same_peak], e64.correlation[same_peak], rtol=1e-6)
    # ground truth: outputs are secondary -> reference offsets
    truth = np.hypot(r64.dx[ok] + 3.30, r64.dy[ok] - 1.70)
    print(f"{prec}: RMS vs the known shift {np.sqrt(np.mean(truth ** 2)):.4f} px", flush=True)
    assert np.median(truth) < 0.1


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("cs,r", [(16, 6), (32, 10), (64, 25), (128, 25), (48, 17)])
def test_geometry_vs_oracle(prec, cs, r, monkeypatch):
    """Chip and radius shapes (every band-height instance), against the CPU oracle directly;
    the 16-bit tracker on every shape, small ones included (AR_MP_MIN_MACS=0)."""
    monkeypatch.setenv("AR_MP_MIN_MACS", "0")
    n = 4 * cs + 4 * r + 40
    ref, sec = shifted_pair(n, (2.4, -1.3), seed=cs + r, dtype=np.float32)
    ref = ref * 255
    sec = sec * 255
    h = cs // 2 + r + 3
    xs = np.linspace(h, n - h, 20)   # 400 points
    pts = autorift.gridpoints_xy(xs, xs, chip_size=cs, search_radius=r)
    print(f"shape c{cs} r{r} {prec}", flush=True)
    p = autorift.params(quantize="none", precision=prec, preprocess="none")
    with autorift.Context(p) as ctx:
        ctx.set_images(ref, sec)
        got = ctx.track(pts, upsampling=1)
        assert ctx.last_kernel()[1] == 2
        pr, ps, pm, _ = ctx.get_pair()
    ok = np.ones(pr.shape, bool)
    exp = oracle.track(pr, ps, ok, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                       pts.chip_size_y, up=1)
    fin = np.isfinite(exp[0])
    assert np.array_equal(np.isfinite(got.dx), fin)
    agree = (got.dx[fin] == exp[0][fin]) & (got.dy[fin] == exp[1][fin])
    print(f"agreement c{cs} r{r} {prec}: {agree.mean():.4f} of {fin.sum()}", flush=True)
    # measured on this texture: 1 of 400 integer peaks differs at chip 16 r 6 in BF16, none
    # elsewhere (such small passes take the exact tracker by default)
    assert agree.mean() >= 0.995, agree.mean()
    # the reported correlation is the exact Float64 value of the chosen peak: where the peak
    # agrees it equals the oracle's to the last bit of the normalisation's box sums
    np.testing.assert_allclose(got.correlation[fin][agree], exp[2][fin][agree], rtol=1e-6)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_ncc_and_padding(prec, monkeypatch):
    """NCC measure, windows past the image edge (virtual zero padding) and a flat chip."""
    n = 160
    ref, sec = shifted_pair(n, (1.0, 2.0), seed=7, dtype=np.float32)
    ref = ref * 200 + 10
    sec = sec * 200 + 10
    sec[60:100, 60:100] = 42.0    # flat: degenerate chip, NaN result but searched
    monkeypatch.setenv("AR_MP_MIN_MACS", "0")   # the 16-bit tracker at chip 32, r 12 too
    xs = np.array([12.0, 40.0, 80.0, 120.0, 150.0])
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=12)
    for measure in (0, 1):
        print(f"ncc/padding {prec} measure {measure}", flush=True)
        got, path = _run(prec, ref, sec, pts, 1, measure=measure, preprocess="none")
        exp, _ = _run("exact", ref, sec, pts, 1, measure=measure, preprocess="none")
        assert path == 2
        rate, _ = _agreement(got, exp)
        assert rate >= 0.95
        assert np.array_equal(np.isnan(got.dx), np.isnan(exp.dx))
    centre = (pts.x == 80.0) & (pts.y == 80.0)
    assert np.isnan(got.dx[centre]).all() and got.searched[centre].all()


def test_fp16_operand_scaling():
    """Values far beyond FP16's range are scaled by a power of two; peaks still agree."""
    n = 200
    ref, sec = shifted_pair(n, (3.0, -2.0), seed=11, dtype=np.float32)
    ref = (ref * 4e6).astype(np.float32)
    sec = (sec * 4e6).astype(np.float32)
    xs = np.arange(50.0, 151.0, 25.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=10)
    print("fp16 scaling", flush=True)
    got, _ = _run("fp16", ref, sec, pts, 1, preprocess="none")
    exp, _ = _run("exact", ref, sec, pts, 1, preprocess="none")
    rate, both = _agreement(got, exp)
    assert both.sum() == len(pts) and rate == 1.0
    assert np.all(np.isfinite(got.correlation))


def test_mixed_chip_sizes_rejected():
    """The 16-bit tracker carves its LDS for one chip size per pass (radius 25: above the
    crossover where small passes take the exact tracker, which accepts mixed sizes)."""
    t = (synthetic_texture((160, 160), seed=3) * 255).astype(np.float32)
    xs = np.array([60.0, 90.0])
    pts = autorift.gridpoints_xy(xs, xs, chip_size=16, search_radius=25)
    pts.chip_size_x[0, 0] = 32
    pts.chip_size_y[0, 0] = 32
    p = autorift.params(quantize="none", precision="bf16")
    with autorift.Context(p) as ctx:
        ctx.set_images(t, t)
        with pytest.raises(autorift.AutoRIFTError, match="one chip size"):
            ctx.track(pts, upsampling=1)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_gaps_and_masks(prec):
    """NaN gaps in the raw images (the pre-filter zeroes them and marks them invalid) and an
    explicit mask: NaN placement and searched flags equal the exact path's, peaks agree."""
    a, b = cfg1_pair(2, n=320)
    a = a.copy()
    a[100:140, 50:200] = np.nan
    m = np.ones(a.shape, bool)
    m[:, 260:] = False
    xs = np.arange(40.0, 281.0, 16.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=12)
    out = []
    for pr in (prec, "exact"):
        with autorift.Context(autorift.params(quantize="none", precision=pr)) as ctx:
            ctx.set_images(a, b, m, m)
            out.append(ctx.track(pts, upsampling=16))
    got, exp = out
    print(f"gaps {prec}", flush=True)
    np.testing.assert_array_equal(got.searched, exp.searched)
    np.testing.assert_array_equal(np.isnan(got.dx), np.isnan(exp.dx))
    fin = np.isfinite(exp.dx)
    err = np.hypot(got.dx[fin] - exp.dx[fin], got.dy[fin] - exp.dy[fin])
    assert np.mean(err == 0) >= 0.98 and (~got.searched).any()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_small_passes_take_the_exact_tracker(prec):
    """Below the measured crossover (8e5 multiply-adds per point) the 16-bit precisions take the
    exact direct tracker, faster there: bit-exact against the oracle; above it, the 16-bit one."""
    n = 200
    ref, sec = shifted_pair(n, (2.4, -1.3), seed=3, dtype=np.float32)
    ref, sec = ref * 255, sec * 255
    xs = np.linspace(50, 150, 6)
    for cs, r, path in ((16, 25, 3), (24, 12, 3), (32, 16, 2)):
        pts = autorift.gridpoints_xy(xs, xs, chip_size=cs, search_radius=r)
        print(f"dispatch {prec} c{cs} r{r}", flush=True)
        p = autorift.params(quantize="none", precision=prec, preprocess="none")
        with autorift.Context(p) as ctx:
            ctx.set_images(ref, sec)
            got = ctx.track(pts, upsampling=64)
            assert ctx.last_kernel()[1] == path
            pr, ps, _, _ = ctx.get_pair()
        if path == 3:
            ok = np.ones(pr.shape, bool)
            exp = oracle.track(pr, ps, ok, pts.x, pts.y, pts.radius_x, pts.radius_y,
                               pts.chip_size_x, pts.chip_size_y, up=64)
            for g, e in zip((got.dx, got.dy, got.correlation), exp[:3]):
                assert np.array_equal(g, e, equal_nan=True)


@pytest.mark.parametrize("prec", ["bf16", "fft"])
def test_few_points_on_a_large_reference_take_the_exact_tracker(prec):
    """The 16-bit and transform trackers build Float64 box images over the whole padded
    reference; on references of at least 4 M chip positions a pass must have the work to repay
    them (searched points x multiply-adds >= 512 x positions), else it takes the exact direct
    tracker. 2100^2 at chip 32, r 25: 9 points fall short, 1681 clear it."""
    n = 2100
    ref, sec = shifted_pair(n, (1.6, -2.2), seed=9, dtype=np.float32)
    ref, sec = ref * 255, sec * 255
    p = autorift.params(quantize="none", precision=prec, preprocess="none")
    with autorift.Context(p) as ctx:
        ctx.set_images(ref, sec)
        for xs, path in ((np.linspace(200, 1900, 3), 3),
                         (np.arange(50.0, 2050.0, 50.0)[:41], 2 if prec == "bf16" else 4)):
            pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
            print(f"large-reference dispatch {prec}: {len(pts)} points", flush=True)
            got = ctx.track(pts, upsampling=16)
            assert ctx.last_kernel()[1] == path
            ok = np.isfinite(got.dx)
            assert ok.mean() > 0.9
            # the tracker reports the secondary's offset as seen from the reference: (-1.6, 2.2)
            assert np.abs(np.median(got.dx[ok]) + 1.6) < 0.1
            assert np.abs(np.median(got.dy[ok]) - 2.2) < 0.1
