"""GPU parity of the per-point tracker against the CPU oracle, through the C ABI.

The bar is bit-exactness: integer peaks, subpixel offsets, correlation values, searched flags
and NaN placement must equal the oracle's (which follows the reference's FP64 arithmetic).
"""
import numpy as np
import pytest

import oracle
from _util import chip_and_window, fixture, shifted_pair, synthetic_texture

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


@pytest.fixture(scope="module")
def ctx():
    c = autorift.Context()
    yield c
    c.close()


def _oracle_point(search, chip, rx, ry, up, measure):
    s, ok = oracle.correlate(search, chip, rx, ry, measure=measure)
    if not ok:
        return np.nan, np.nan, np.nan
    if up == 1:
        i, j = oracle.peak_index(s)
        return float(j - rx - 1), float(i - ry - 1), float(s[i - 1, j - 1])
    dx, dy, v = oracle.subpixel_peak(s, rx, ry, up)
    return dx, dy, v


def _same(a, b):
    a32, b32 = np.float32(a), np.float32(b)
    return (np.isnan(a32) and np.isnan(b32)) or a32.tobytes() == b32.tobytes()


ZNCC = [(c, r, d, v) for c in (32, 64, 128) for r in (6, 10, 25) for d in ("uint8", "float32")
        for v in ("plain",)] + [(c, r, "float32", "dc_offset") for c in (32, 64) for r in (6, 25)]


@pytest.mark.parametrize("cs,r,dt,variant", ZNCC)
@pytest.mark.parametrize("up", [1, 64])
def test_correlate_point_fixtures(ctx, cs, r, dt, variant, up):
    a, _ = fixture(f"matchtemplate/zncc_c{cs}_r{r}_{dt}_{variant}")
    print(f"case c{cs} r{r} {dt} {variant} up{up}", flush=True)
    got = ctx.correlate_point(a["search"], a["chip"], r, upsampling=up)
    exp = _oracle_point(a["search"], a["chip"], r, r, up, oracle.ZNCC)
    for g, e in zip(got, exp):
        assert _same(g, e), (got, exp)
    # the OpenCV fixture's own integer peak
    if up == 1:
        i, j = oracle.peak_index(a["expected"])
        assert (got[0], got[1]) == (j - r - 1, i - r - 1)


@pytest.mark.parametrize("cs,r", [(32, 25), (16, 6), (64, 10)])
@pytest.mark.parametrize("dt", [np.uint8, np.float32])
def test_lowcontrast_and_measures(ctx, cs, r, dt):
    rng = np.random.default_rng(cs + r)
    search = (rng.random((cs + 2 * r - 1, cs + 2 * r - 1)) * 255).astype(dt)
    chip = search[r:r + cs, r:r + cs].copy()
    for measure in (oracle.ZNCC, oracle.NCC):
        for up in (1, 2, 16, 32, 64):
            print(f"case c{cs} r{r} {np.dtype(dt)} m{measure} up{up}", flush=True)
            got = ctx.correlate_point(search, chip, r, upsampling=up, measure=measure)
            exp = _oracle_point(search, chip, r, r, up, measure)
            for g, e in zip(got, exp):
                assert _same(g, e), (measure, up, got, exp)


def _scene(n, dtype, seed, shift=(2.5, -1.25)):
    ref, sec = shifted_pair(n, shift, seed=seed, dtype=dtype)
    return ref, sec


def _compare(ctx, ref, sec, ok, pts, up, measure=oracle.ZNCC):
    ctx.set_pair(ref, sec, ok, ok)
    print(f"track n={len(pts)} up={up} m={measure} {ref.dtype}", flush=True)
    g = ctx.track(pts, upsampling=up, measure=measure)
    e = oracle.track(ref, sec, ok, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                     pts.chip_size_y, pts.dx_prior, pts.dy_prior, up=up, measure=measure,
                     nthreads=8)
    np.testing.assert_array_equal(g.searched, e[3])
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(a.view(np.uint32)[~np.isnan(a)],
                                      b.view(np.uint32)[~np.isnan(b)])
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    return g


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("up", [1, 64])
def test_track_grid_with_padding(ctx, dtype, up):
    ref, sec = _scene(160, dtype, seed=3)
    ok = np.ones(ref.shape, bool)
    # 12-px border: windows cross the image edge, so the pass pads (src/track.jl:128-138)
    xs = np.arange(12, 150, 11, dtype=float)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    g = _compare(ctx, ref, sec, ok, pts, up)
    assert g.searched.all()


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_track_mixed_geometry(ctx, dtype):
    rng = np.random.default_rng(7)
    ref, sec = _scene(192, dtype, seed=4, shift=(-3.3, 1.7))
    ok = np.ones(ref.shape, bool)
    ok[40:70, 100:140] = False           # masked patch
    ref = ref.copy(); sec = sec.copy()
    ref[150:, :40] = ref[150, 0]         # flat corner: degenerate chips
    sec[150:, :40] = sec[150, 0]
    n = 300
    x = rng.uniform(-20, 212, n)
    y = rng.uniform(-20, 212, n)
    cs = rng.choice([16, 32, 64], n)
    rx = rng.integers(0, 20, n)
    ry = rng.integers(0, 20, n)
    pts = autorift.pointset(x, y, search_radius_x=rx, search_radius_y=ry, chip_size_x=cs,
                            chip_size_y=cs, dx_prior=rng.normal(0, 3, n),
                            dy_prior=rng.normal(0, 3, n))
    for up in (1, 16, 64):
        _compare(ctx, ref, sec, ok, pts, up)
    _compare(ctx, ref, sec, ok, pts, 64, measure=oracle.NCC)


def test_track_nonpow2_chip_u8_uses_direct_f64(ctx):
    ref, sec = _scene(160, np.uint8, seed=5)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(60, 110, 16, dtype=float)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=48, search_radius=12)
    _compare(ctx, ref, sec, ok, pts, 64)
    assert ctx.last_kernel()[1] == 3
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=12)
    _compare(ctx, ref, sec, ok, pts, 64)
    assert ctx.last_kernel()[1] == 0


def test_track_known_shift_recovery(ctx):
    # ground truth: integer shift recovered exactly everywhere (test/track.jl:15-52)
    ref, sec = shifted_pair(256, (4, -3), seed=9, dtype=np.uint8)
    ok = np.ones(ref.shape, bool)
    xs = np.arange(60, 200, 16, dtype=float)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    ctx.set_pair(ref, sec, ok, ok)
    g = ctx.track(pts, upsampling=1)
    assert np.all(g.dx == -4) and np.all(g.dy == 3)
    g = ctx.track(pts, upsampling=64)
    assert np.all(np.abs(g.dx + 4) <= 3 / 64) and np.all(np.abs(g.dy - 3) <= 3 / 64)


def _energy_in_order(chip):
    """sum (x - mean)^2 of a chip in the reference's order (src/correlate.jl:435-447), as plain
    Python float additions, which are sequential."""
    v = [float(x) for x in chip.ravel(order="F")]
    s = 0.0
    for x in v:
        s += x
    mean = s / len(v)
    sq = 0.0
    for x in v:
        d = x - mean
        sq += d * d
    return sq


def _energy_exact(chip):
    """The same sum without rounding, rounded once."""
    n = chip.size
    s = int(chip.sum(dtype=np.int64))
    s2 = int((chip.astype(np.int64) ** 2).sum())
    return float((n * n * s2 - n * s * s)) / n / n


@pytest.mark.parametrize("odd", [False, True])
def test_u8_chip_energy_in_reference_order(ctx, odd):
    """128 x 128 UInt8 chips of high contrast: past 2^53 / n^2 the reference's sequential energy
    sum rounds (every step ties when the chip sum is odd), so an exact integer energy differs from
    it in the last bits. The device must take the reference's; the case is checked to have that
    difference, and the correlation surface's peak and value are compared bit for bit."""
    rng = np.random.default_rng(17 + odd)
    r = 8
    found = 0
    for k in range(6):
        base = rng.integers(0, 2, (128 + 2 * r + 15, 128 + 2 * r + 15)) * 255
        base = np.clip(base + rng.integers(-40, 41, base.shape), 0, 255).astype(np.uint8)
        search = np.ascontiguousarray(base[5:5 + 128 + 2 * r - 1, 7:7 + 128 + 2 * r - 1])
        chip = base[5 + r + 2:5 + r + 2 + 128, 7 + r - 3:7 + r - 3 + 128].copy()
        if (int(chip.sum(dtype=np.int64)) % 2 == 1) != odd:
            chip[0, 0] = chip[0, 0] + 1 if chip[0, 0] < 255 else chip[0, 0] - 1
        if _energy_in_order(chip) != _energy_exact(chip):
            found += 1
        print(f"u8 energy c128 r{r} odd={odd} case {k}", flush=True)
        for up in (1, 64):
            got = ctx.correlate_point(search, chip, r, upsampling=up)
            assert ctx.last_kernel()[1] == 0   # the int8 matrix-core tracker
            exp = _oracle_point(search, chip, r, r, up, oracle.ZNCC)
            for g, e in zip(got, exp):
                assert _same(g, e), (k, up, got, exp)
    assert found >= 3   # the cases exercise the rounding sequential sum
