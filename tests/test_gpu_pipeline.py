"""GPU parity of image preparation and the multi-chip-size search against the CPU oracle."""
import numpy as np
import pytest

import oracle
from _util import shifted_pair, synthetic_texture

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


def _params(**kw):
    p = autorift.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_prepare(img, mask, quantize=True, w=5):
    hp, vm = oracle.highpass_filtered(img, mask, w)
    if quantize:
        return oracle.to_uint8(hp, vm), vm
    return hp, vm   # already finite: _filtered zeroes non-finite values


def _images(kind, n=(96, 130), seed=0):
    rng = np.random.default_rng(seed)
    t = synthetic_texture(n, seed=seed)
    if kind == "u8":
        return np.rint(t * 255).astype(np.uint8), None
    if kind == "i16":
        return np.rint(t * 4000 - 2000).astype(np.int16), None
    if kind == "f64":
        return (t * 100).astype(np.float64), None
    img = (t * 255).astype(np.float32)
    if kind == "f32_nan":
        img[rng.random(n) < 0.01] = np.nan
        img[10:20, 30:45] = np.nan
        return img, None
    if kind == "f32_mask":
        m = np.ones(n, bool)
        m[:, :7] = False
        m[50:58, 60:90] = False
        return img, m
    if kind == "f32_tiny":
        # values 2^-40 apart from their neighbours' scale: the box sums round, so the
        # preparation must take the reference's sequential running sums
        img[rng.random(n) < 0.05] = np.float32(3e-12)
        img[40:44, 20:30] *= np.float32(1 + 2 ** -20)
        return img, None
    if kind == "f32_inf":
        img[30, 40] = np.inf
        img[70, 100] = -np.inf
        return img, None
    return img, None


@pytest.mark.parametrize("kind", ["u8", "i16", "f32", "f64", "f32_nan", "f32_mask", "f32_tiny",
                                  "f32_inf"])
@pytest.mark.parametrize("quant", [True, False])
def test_prepare_matches_oracle(kind, quant):
    img, mask = _images(kind)
    img2, _ = _images(kind, seed=1)
    p = _params(quantize=1 if quant else 0)
    print(f"prepare {kind} quantize={quant}", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_images(img, img2, mask, mask)
        r, s, rv, sv = ctx.get_pair()
    m = np.isfinite(img.astype(np.float64)) if mask is None else mask
    er, evm = _oracle_prepare(img, m, quant)
    np.testing.assert_array_equal(rv, evm)
    if quant:
        assert r.dtype == np.uint8
        np.testing.assert_array_equal(r, er)
    else:
        assert r.dtype == np.float32
        np.testing.assert_array_equal(r.view(np.uint32), np.asarray(er, np.float32).view(np.uint32))


@pytest.mark.parametrize("kind", ["u8", "i16", "f32", "f64", "f32_nan", "f32_mask"])
@pytest.mark.parametrize("w,min_std", [(5, 0.0), (7, 2.0), (21, 0.0)])
def test_wallis_prepare_matches_oracle(kind, w, min_std):
    """Wallis (src/preprocess.jl:158-159, 333-351, 416-438) on the device, bit-exact against the
    oracle's restatement, both quantisers."""
    img, mask = _images(kind)
    img2, _ = _images(kind, seed=1)
    m = np.isfinite(img.astype(np.float64)) if mask is None else mask
    er, evm = oracle.wallis_filtered(img, m, w, min_std)
    print(f"wallis {kind} w {w} min_std {min_std}", flush=True)
    for quant in (True, False):
        p = _params(quantize=1 if quant else 0, preprocess=3, filter_width=w, min_std=min_std)
        with autorift.Context(p) as ctx:
            ctx.set_images(img, img2, mask, mask)
            r, _, rv, _ = ctx.get_pair()
        np.testing.assert_array_equal(rv, evm)
        if quant:
            np.testing.assert_array_equal(r, oracle.to_uint8(er, evm))
        else:
            np.testing.assert_array_equal(r.view(np.uint32), er.view(np.uint32))
    # (scattered NaNs erode most of the image at width 21: a pixel needs its whole window valid)
    assert evm.sum() > (0 if kind == "f32_nan" else evm.size // 2)


def test_wallis_multichip_matches_oracle():
    """The default search over a Wallis-prepared pair (width 7): device preparation and the
    multi-chip pass, bit-exact against the oracle's Wallis + quantiser + multichip."""
    a, b = shifted_pair(320, (2.25, -1.5), seed=8)
    a, b = a * np.float32(255), b * np.float32(255)
    p = autorift.params(preprocess=autorift.Wallis(width=7, min_std=0.5), chip_size=32)
    xs = np.arange(72.0, 249.0, 16.0)
    grid = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    print("wallis multichip", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_images(a, b)
        got = ctx.run_multichip(grid)
    ok = np.isfinite(a)
    er, erv = oracle.wallis_filtered(a, ok, 7, 0.5)
    es, esv = oracle.wallis_filtered(b, np.isfinite(b), 7, 0.5)
    e = oracle.multichip(oracle.to_uint8(er, erv), oracle.to_uint8(es, esv), erv & esv, grid.x,
                         grid.y, grid.radius_x, grid.radius_y, p=oracle.params(chip_size=32),
                         nthreads=16)
    for g, x in zip((got.dx, got.dy, got.correlation), e[:3]):
        np.testing.assert_array_equal(np.asarray(g, np.float32).view(np.uint32),
                                      np.asarray(x, np.float32).view(np.uint32))
    np.testing.assert_array_equal(got.chip_size, e[3])
    assert np.isfinite(got.dx).mean() > 0.9


def test_wallis_rejects_complex_input():
    z = (np.ones((64, 64)) + 1j * np.ones((64, 64))).astype(np.complex64)
    p = autorift.params(similarity="coherence", preprocess="wallis", quantize="none")
    with autorift.Context(p) as ctx:
        with pytest.raises(autorift.ArgumentError, match="Wallis"):
            ctx.set_images(z, z)


def test_prepare_reinit_keeps_other_slot():
    a, _ = _images("f32", seed=2)
    b, _ = _images("f32", seed=3)
    c, _ = _images("f32", seed=4)
    with autorift.Context() as ctx:
        ctx.set_images(a, b)
        r0, s0, _, _ = ctx.get_pair()
        ctx.set_images(secondary=c)
        r1, s1, _, _ = ctx.get_pair()
        np.testing.assert_array_equal(r0, r1)
        ec, _ = _oracle_prepare(c, np.ones(c.shape, bool))
        np.testing.assert_array_equal(s1, ec)
        with pytest.raises(autorift.DimensionMismatch):
            ctx.set_images(secondary=c[:50])


def _grid(n, spacing, margin):
    xs = np.arange(margin + 1, n - margin + 1, spacing, dtype=np.float64)
    return autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)


@pytest.mark.parametrize("case", ["shift_u8", "shift_f32_masked", "warp_u8_levels", "noise_gate",
                                  "all_masked"])
def test_multichip_matches_oracle(case):
    if case == "noise_gate":
        # uncorrelated images and a strict coherence gate: levels end in the coarse pass
        rng = np.random.default_rng(11)
        ref = rng.integers(0, 256, (300, 300)).astype(np.uint8)
        sec = rng.integers(0, 256, (300, 300)).astype(np.uint8)
        ok = np.ones(ref.shape, bool)
        grid = _grid(300, 6, 60)
        p = _params(preprocess=0, quantize=0, chip_size_base=16, chip_size_min=16,
                    chip_size_max=64, min_coarse_valid_fraction=0.5)
        op = oracle.params(chip_size=16, chip_size_max=64, min_coarse_valid_fraction=0.5)
    elif case == "all_masked":
        ref, sec = shifted_pair(300, (1.0, 1.0), seed=9, dtype=np.uint8)
        ok = np.zeros(ref.shape, bool)
        grid = _grid(300, 6, 60)
        p = _params(preprocess=0, quantize=0)
        op = oracle.params(chip_size=32)
    elif case == "shift_u8":
        ref, sec = shifted_pair(400, (4.4, -3.2), seed=5, dtype=np.uint8)
        ok = np.ones(ref.shape, bool)
        grid = _grid(400, 8, 90)
        p = _params(preprocess=0, quantize=0)
        op = oracle.params(chip_size=32)
    elif case == "shift_f32_masked":
        ref, sec = shifted_pair(300, (-2.6, 1.9), seed=6, dtype=np.float32)
        ok = np.ones(ref.shape, bool)
        ok[100:140, 60:120] = False
        grid = _grid(300, 6, 60)
        p = _params(preprocess=0, quantize=0, chip_size_base=16, chip_size_min=16,
                    chip_size_max=64)
        op = oracle.params(chip_size=16, chip_size_max=64)
    else:
        # low texture in one half forces coarser levels to take over
        ref, sec = shifted_pair(360, (1.5, 2.5), seed=7, dtype=np.float32)
        ref = ref * 255
        sec = sec * 255
        ref[:, 180:] = ref[:, 180:] * 0.02 + 100
        sec[:, 180:] = sec[:, 180:] * 0.02 + 100
        rng = np.random.default_rng(1)
        sec = (sec + rng.normal(0, 1, sec.shape)).astype(np.float32)
        ref = np.clip(np.rint(ref), 0, 255).astype(np.uint8)
        sec = np.clip(np.rint(sec), 0, 255).astype(np.uint8)
        ok = np.ones(ref.shape, bool)
        grid = _grid(360, 7, 60)
        p = _params(preprocess=0, quantize=0, chip_size_base=16, chip_size_min=16,
                    chip_size_max=64)
        op = oracle.params(chip_size=16, chip_size_max=64)
    print(f"multichip {case} grid {grid.shape}", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        g = ctx.run_multichip(grid)
    e = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior, p=op, nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])
    np.testing.assert_array_equal(g.interpolated, e[4])
    if case == "all_masked":
        assert g.nmeasured() == 0
    elif case == "noise_gate":
        print(f"noise_gate: {g.nmeasured()} of {g.dx.size} measured", flush=True)
    else:
        assert g.nmeasured() > 0.5 * g.dx.size


def test_multichip_refuses_more_measures_than_levels():
    """_check_measures (src/types.jl:899-906) on the device path: a caller of the C ABI whose
    similarity tuple is longer than the chip-size levels gets the reference's ArgumentError."""
    p = _params(chip_size_base=32, chip_size_min=32, chip_size_max=64, preprocess=0, quantize=0)
    p.level_measure_count = 3
    for k in range(3):
        p.level_measures[k] = 0
    ref, sec = shifted_pair(200, (1, 1), seed=3, dtype=np.uint8)
    xs = np.arange(80.0, 121.0, 20.0)
    grid = autorift.gridpoints_xy(xs, xs, chip_size=64, search_radius=10)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec)
        with pytest.raises(autorift.ArgumentError, match="names 3 measures but there are only 2"):
            ctx.run_multichip(grid)


@pytest.mark.parametrize("buffer", [1, 2, 30])
def test_multichip_coarse_dilation_radii(buffer):
    """The coarse mask's dilation (dilate_within, src/resample.jl:211-219) at radii the direct
    disc form takes (1, 2) and one the distance-transform form takes (30), on a grid whose coarse
    mask has holes (a low-texture half): bit-exact against the oracle's transform."""
    ref, sec = shifted_pair(360, (1.5, 2.5), seed=7, dtype=np.float32)
    ref, sec = ref * 255, sec * 255
    ref[:, 180:] = ref[:, 180:] * 0.02 + 100
    sec[:, 180:] = sec[:, 180:] * 0.02 + 100
    rng = np.random.default_rng(2)
    sec = (sec + rng.normal(0, 1, sec.shape)).astype(np.float32)
    ref = np.clip(np.rint(ref), 0, 255).astype(np.uint8)
    sec = np.clip(np.rint(sec), 0, 255).astype(np.uint8)
    ok = np.ones(ref.shape, bool)
    grid = _grid(360, 5, 55)
    p = _params(preprocess=0, quantize=0, chip_size_base=16, chip_size_min=16, chip_size_max=64,
                coarse_buffer=buffer)
    op = oracle.params(chip_size=16, chip_size_max=64, coarse_buffer=buffer)
    print(f"multichip coarse_buffer {buffer} grid {grid.shape}", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        g = ctx.run_multichip(grid)
    e = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior, p=op, nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])
    np.testing.assert_array_equal(g.interpolated, e[4])
    print(f"  measured {g.nmeasured()} of {g.dx.size}, levels "
          f"{dict(zip(*np.unique(g.chip_size, return_counts=True)))}", flush=True)


def test_multichip_far_points_match_oracle():
    """Grid points far outside the image: one beyond 1e15 (the host then plans no level geometry
    up front and every pass takes its padding from the device) and one thousands of pixels off
    (unsearched; hole filling may give it its neighbours' value): bit-exact against the oracle."""
    ref, sec = shifted_pair(300, (2.0, -1.5), seed=8, dtype=np.uint8)
    ok = np.ones(ref.shape, bool)
    grid = _grid(300, 12, 60)
    grid.x[0, 0] = 2e15
    grid.y[1, 1] = -5000.0
    p = _params(preprocess=0, quantize=0)
    op = oracle.params(chip_size=32)
    print(f"multichip far points grid {grid.shape}", flush=True)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        g = ctx.run_multichip(grid)
    e = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior, p=op, nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])
    np.testing.assert_array_equal(g.interpolated, e[4])
    assert g.nmeasured() > 0.5 * g.dx.size
