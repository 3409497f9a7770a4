"""Pins the CPU oracle against the reference's golden fixtures, closed forms and brute force.

Every function of the oracle that the GPU path is compared against is checked here by
something that does not retype its formula: OpenCV fixtures (matchTemplate, pyrUp,
minMaxLoc, distanceTransform), literal brute-force reductions, and known-answer scenes.
"""
import itertools
import os

import numpy as np
import pytest

import oracle
from _util import GOLDEN, chip_and_window, fixture, shifted_pair, synthetic_texture


# --------------------------------------------------------------------------------------------
# correlation surface vs OpenCV matchTemplate(TM_CCOEFF_NORMED) -- test/opencv.jl:22-68
# --------------------------------------------------------------------------------------------

def _zncc_cases():
    out = []
    for cs, r, dt, variant in itertools.product((32, 64, 128), (6, 10, 25), ("uint8", "float32"),
                                                 ("plain", "dc_offset")):
        name = f"matchtemplate/zncc_c{cs}_r{r}_{dt}_{variant}"
        if os.path.isdir(os.path.join(GOLDEN, name)):
            out.append((name, r))
    return out


ZNCC_CASES = _zncc_cases()


def test_fixture_count():
    # 3 chip sizes x 3 radii x (2 dtypes plain + float32 dc_offset) = 27 (test/opencv.jl:59)
    assert len(ZNCC_CASES) == 27


@pytest.mark.parametrize("mode", [oracle.NUM_DIRECT, oracle.NUM_FFT])
@pytest.mark.parametrize("name,r", ZNCC_CASES)
def test_surface_vs_opencv(name, r, mode):
    a, _ = fixture(name)
    got, ok = oracle.correlate(a["search"], a["chip"], r, num_mode=mode)
    assert ok
    exp = a["expected"]
    assert got.shape == exp.shape
    assert oracle.peak_index(got) == oracle.peak_index(exp)
    assert np.max(np.abs(got.astype(np.float64) - exp)) < 1e-4


@pytest.mark.parametrize("cs,r", [(c, r) for c in (32, 64, 128) for r in (6, 10, 25)])
def test_lowcontrast_truth(cs, r):
    # test/opencv.jl:88-139: OpenCV gives zeros; Float64 integral images recover the centre.
    a, _ = fixture(f"matchtemplate/zncc_c{cs}_r{r}_float32_lowcontrast")
    assert np.all(a["expected"] == 0)
    got, ok = oracle.correlate(a["search"], a["chip"], r)
    assert oracle.peak_index(got) == (r + 1, r + 1)
    assert abs(got.max() - 1.0) < 1e-3
    assert np.all(np.isfinite(got))
    bound = np.float32(1.0) + 2048 * np.finfo(np.float32).eps
    assert np.all((-bound <= got) & (got <= bound))


def test_dc_offset_cancels():
    p, _ = fixture("matchtemplate/zncc_c32_r25_float32_plain")
    o, _ = fixture("matchtemplate/zncc_c32_r25_float32_dc_offset")
    s1, _ = oracle.correlate(p["search"], p["chip"], 25)
    s2, _ = oracle.correlate(o["search"], o["chip"], 25)
    assert oracle.peak_index(s1) == oracle.peak_index(s2)
    assert np.max(np.abs(s1 - s2)) < 1e-4


def _brute_zncc(search, chip, rx, ry):
    """Literal per-shift Pearson correlation in float64 (no integral images, no mean trick)."""
    ch, cw = chip.shape
    t = chip.astype(np.float64) - chip.mean(dtype=np.float64)
    out = np.zeros((2 * ry, 2 * rx))
    for i in range(2 * ry):
        for j in range(2 * rx):
            w = search[i:i + ch, j:j + cw].astype(np.float64)
            w = w - w.mean()
            den = np.sqrt((t * t).sum() * (w * w).sum())
            out[i, j] = (t * w).sum() / den if den > 0 else 0.0
    return out


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_surface_vs_brute_force_nonsquare(dtype):
    rng = np.random.default_rng(11)
    ch, cw, rx, ry = 12, 20, 5, 3
    search = (rng.random((ch + 2 * ry - 1, cw + 2 * rx - 1)) * 255).astype(dtype)
    chip = (rng.random((ch, cw)) * 255).astype(dtype)
    got, ok = oracle.correlate(search, chip, rx, ry)
    assert ok and got.shape == (2 * ry, 2 * rx)
    np.testing.assert_allclose(got, _brute_zncc(search, chip, rx, ry), atol=2e-6)


def test_ncc_closed_form():
    # NCC (src/correlate.jl:604-630): mean-removed chip over the raw window energy.
    rng = np.random.default_rng(5)
    ch = cw = 8
    r = 4
    search = rng.random((ch + 2 * r - 1, cw + 2 * r - 1)).astype(np.float32) + 1
    chip = rng.random((ch, cw)).astype(np.float32)
    got, _ = oracle.correlate(search, chip, r, measure=oracle.NCC)
    t = chip.astype(np.float64) - chip.astype(np.float64).mean()
    exp = np.zeros_like(got, dtype=np.float64)
    for i in range(2 * r):
        for j in range(2 * r):
            w = search[i:i + ch, j:j + cw].astype(np.float64)
            exp[i, j] = (t * w).sum() / np.sqrt((t * t).sum() * (w * w).sum())
    np.testing.assert_allclose(got, exp, atol=2e-6)


def test_exact_recovery_and_self_correlation():
    img = synthetic_texture((300, 300), seed=7).astype(np.float32)
    for sx, sy in ((0, 0), (3, -2), (7, 11), (-13, 5), (24, -24), (-24, 24)):
        chip, search = chip_and_window(img, (150, 150), 32, 25, shift=(sy, sx))
        s, ok = oracle.correlate(search, chip, 25)
        i, j = oracle.peak_index(s)
        assert (j - 25 - 1, i - 25 - 1) == (sx, sy)
        assert s[i - 1, j - 1] > 0.999


def test_degenerate_chip():
    search = synthetic_texture((31, 31), seed=1).astype(np.float32)
    s, ok = oracle.correlate(search, np.full((16, 16), 0.5, np.float32), 8)
    assert not ok and np.all(s == 0)


def test_integral_closed_forms():
    A = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float32)
    S, S2 = oracle.integral_both(A)
    assert np.all(S[0] == 0) and np.all(S[:, 0] == 0)
    assert S[-1, -1] == A.sum() and S2[-1, -1] == (A.astype(np.float64) ** 2).sum()
    np.testing.assert_array_equal(S[1:, 1:], A.astype(np.float64).cumsum(0).cumsum(1))
    B = np.full((64, 64), 255, np.uint8)
    assert oracle.integral_both(B)[1][-1, -1] == 64 * 64 * 255.0 ** 2


def test_next_fft_size():
    assert oracle.next_fft_size(1) == 1
    assert oracle.next_fft_size(43) == 48
    assert oracle.next_fft_size(113) == 128
    assert 81 < oracle.next_fft_size(81) < 128
    assert oracle.next_fft_size(81) == 84
    for n in (17, 31, 43, 81, 83, 113, 127, 163, 177, 251):
        m = oracle.next_fft_size(n)
        assert n <= m <= int(np.ceil(1.5 * n))
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        assert r == 1


# --------------------------------------------------------------------------------------------
# peak and subpixel
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["unique", "tie_in_row", "tie_in_column", "plateau", "all_equal"])
def test_peak_vs_minmaxloc(name):
    a, p = fixture(f"peak/{name}")
    r, c = p["row_col_0based"]
    assert oracle.peak_index(a["surface"]) == (r + 1, c + 1)


def test_peak_hand_cases():
    a = np.zeros((7, 7), np.float32); a[2, 5] = 1
    assert oracle.peak_index(a) == (3, 6)
    a = np.zeros((7, 7), np.float32); a[3, 1] = a[3, 5] = 1
    assert oracle.peak_index(a) == (4, 2)
    a = np.zeros((7, 7), np.float32); a[1, 3] = a[5, 3] = 1
    assert oracle.peak_index(a) == (2, 4)
    a = np.zeros((7, 7), np.float32); a[2:5, 2:5] = 1
    assert oracle.peak_index(a) == (3, 3)
    a = np.zeros((5, 5), np.float32); a[1, 1] = np.nan; a[3, 3] = 0.5
    assert oracle.peak_index(a) == (4, 4)
    assert oracle.peak_index(np.full((3, 3), np.nan, np.float32)) == (1, 1)


def _pyrup_cases(kind):
    d = os.path.join(GOLDEN, "pyrup")
    return sorted(c for c in os.listdir(d) if f"_{kind}_" in c)


@pytest.mark.parametrize("name", _pyrup_cases("step"))
def test_pyrup_step_vs_opencv(name):
    a, _ = fixture(f"pyrup/{name}")
    got = oracle.pyrup(a["src"])
    assert np.max(np.abs(got - a["expected"])) < 1e-6


@pytest.mark.parametrize("name", _pyrup_cases("cascade"))
def test_pyrup_cascade_vs_opencv(name):
    a, p = fixture(f"pyrup/{name}")
    cur = a["src"]
    k = 1
    while k < p["factor"]:
        cur = oracle.pyrup(cur)
        k *= 2
    assert np.max(np.abs(cur - a["expected"])) < 1e-4


def test_subpixel_properties():
    r = 25
    a = np.zeros((2 * r, 2 * r), np.float32)
    a[r + 3, r + 7] = 1
    assert oracle.subpixel_peak(a, r, r, 1)[:2] == (7.0, 3.0)
    a = np.zeros((2 * r, 2 * r), np.float32)
    for di in range(-2, 3):
        for dj in range(-2, 3):
            a[r + di, r + dj] = np.exp(-(di * di + dj * dj) / 2)
    dx, dy, _ = oracle.subpixel_peak(a, r, r, 16)
    assert abs(dx) < 0.1 and abs(dy) < 0.1
    a = np.zeros((2 * r, 2 * r), np.float32)
    a[r, r] = 1.0; a[r, r + 1] = 0.9; a[r, r - 1] = 0.3
    assert oracle.subpixel_peak(a, r, r, 16)[0] > 0
    for up in (2, 4, 16, 64):
        dx, dy, _ = oracle.subpixel_peak(a, r, r, up)
        assert abs(dx * up - round(dx * up)) < 1e-9 and abs(dy * up - round(dy * up)) < 1e-9
    # edge clamping and the small-surface fallback (test/peak.jl:184-240)
    r = 10
    a = np.zeros((2 * r, 2 * r), np.float32); a[0, 0] = 1
    dx, dy, v = oracle.subpixel_peak(a, r, r, 16)
    assert v > 0 and abs(dx + r) <= 1 and abs(dy + r) <= 1
    small = np.zeros((4, 4), np.float32); small[1, 2] = 1
    assert oracle.subpixel_peak(small, 2, 2, 16)[:2] == (0.0, -1.0)


def test_subpixel_fractional_recovery():
    cs, r, up = 32, 10, 32
    for sx, sy in ((0.5, 0.0), (0.0, -0.5), (0.25, 0.75), (-1.5, 2.25)):
        ref, sec = shifted_pair(160, (sx, sy), seed=0)
        chip, search = chip_and_window(sec, (80, 80), cs, r, window_from=ref)
        s, _ = oracle.correlate(search, chip, r)
        dx, dy, c = oracle.subpixel_peak(s, r, r, up)
        assert abs(-dx - sx) < 0.1 and abs(-dy - sy) < 0.1 and c > 0.9


# --------------------------------------------------------------------------------------------
# window reductions vs brute force -- test/window.jl:11-48, 111-151
# --------------------------------------------------------------------------------------------

def _brute_window(a, w, reducer):
    nr, nc = a.shape
    wx = min(w, nc); wy = min(w, nr)
    lx, ly = wx // 2, wy // 2
    rx, ry = wx - 1 - lx, wy - 1 - ly
    out = np.full(a.shape, np.nan)
    for i in range(nr):
        for j in range(nc):
            v = a[max(i - ly, 0):min(i + ry, nr - 1) + 1, max(j - lx, 0):min(j + rx, nc - 1) + 1]
            v = v[~np.isnan(v)]
            if v.size:
                out[i, j] = reducer(v)
    return out


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("w", [1, 2, 3, 4, 5, 8])
def test_windows_vs_brute(seed, w):
    rng = np.random.default_rng(seed)
    a = rng.random((rng.integers(3, 40), rng.integers(3, 40))).astype(np.float32)
    a[rng.random(a.shape) < (0.3 if seed % 2 else 0.0)] = np.nan
    np.testing.assert_allclose(oracle.windowmean(a, w), _brute_window(a, w, np.mean),
                               rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(oracle.window_extremum(a, w, False),
                                  _brute_window(a, w, np.min).astype(np.float32))
    np.testing.assert_array_equal(oracle.window_extremum(a, w, True),
                                  _brute_window(a, w, np.max).astype(np.float32))
    med, mad = oracle.windowmedmad(a, w)
    bm = _brute_window(a, w, lambda v: np.median(v.astype(np.float64)))
    bmad = _brute_window(a, w, lambda v: np.median(np.abs(v - np.float32(np.median(v)))))
    np.testing.assert_allclose(med, bm, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(mad, bmad, rtol=1e-5, atol=1e-6)


def test_count_agreeing_brute():
    rng = np.random.default_rng(3)
    a = rng.random((17, 23)).astype(np.float32)
    a[rng.random(a.shape) < 0.2] = np.nan
    got = oracle.count_agreeing(a, 5, 0.2)
    nr, nc = a.shape
    for i in range(nr):
        for j in range(nc):
            c = a[i, j]
            if np.isnan(c):
                assert got[i, j] == 0
                continue
            v = a[max(i - 2, 0):i + 3, max(j - 2, 0):j + 3]
            assert got[i, j] == np.sum(np.abs(v - c) < np.float32(0.2))


# --------------------------------------------------------------------------------------------
# EDT vs OpenCV distanceTransform(DIST_L2, precise) and brute force; nearest resample
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["random", "single", "stripes"])
def test_edt_vs_opencv(name):
    a, _ = fixture(f"disttransform/{name}")
    d2 = oracle.sqdist_transform(a["src"] == 0)
    np.testing.assert_allclose(np.sqrt(d2), a["expected"], atol=1e-4)


def test_dilate_brute():
    rng = np.random.default_rng(2)
    m = rng.random((23, 31)) < 0.05
    got = oracle.dilate_within(m, 8)
    pts = np.argwhere(m)
    I, J = np.meshgrid(np.arange(23), np.arange(31), indexing="ij")
    d2 = np.full(m.shape, np.inf)
    for (i, j) in pts:
        d2 = np.minimum(d2, (I - i) ** 2 + (J - j) ** 2)
    np.testing.assert_array_equal(got, d2 <= 64)
    assert not oracle.dilate_within(np.zeros((5, 5), bool), 8).any()


def test_resample_nearest_closed_form():
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    out = oracle.resample_nearest(a, (12, 16))
    # every 4x4 destination block maps onto its source sample
    np.testing.assert_array_equal(out, np.kron(a, np.ones((4, 4), np.float32)))
    out = oracle.resample_nearest(a, (3, 4))
    np.testing.assert_array_equal(out, a)


# --------------------------------------------------------------------------------------------
# Gardner filter and hole filling -- test/outliers.jl, test/multichip.jl properties
# --------------------------------------------------------------------------------------------

def test_gardner_properties():
    rng = np.random.default_rng(4)
    nr, nc = 30, 30
    dx = np.full((nr, nc), 2.0, np.float32) + rng.normal(0, 0.05, (nr, nc)).astype(np.float32)
    dy = np.full((nr, nc), -1.0, np.float32) + rng.normal(0, 0.05, (nr, nc)).astype(np.float32)
    bad = [(5, 5), (12, 20), (25, 3)]
    for (i, j) in bad:
        dx[i, j] = 20.0
    r = np.full((nr, nc), 25, np.int64)
    valid = np.ones((nr, nc), bool)
    keep = oracle.reject_gardner(dx, dy, r, r, valid, 64)
    for (i, j) in bad:
        assert not keep[i, j]
    assert keep.sum() > 0.9 * nr * nc
    # never resurrects an invalid point; more iterations only remove
    valid2 = valid.copy(); valid2[10, 10] = False
    assert not oracle.reject_gardner(dx, dy, r, r, valid2, 64)[10, 10]
    k1 = oracle.reject_gardner(dx, dy, r, r, valid, 64, oracle.gardner(iterations=1))
    k3 = oracle.reject_gardner(dx, dy, r, r, valid, 64, oracle.gardner(iterations=3))
    assert np.all(k3 <= k1)


def test_fill_holes():
    dx = np.ones((7, 7), np.float32)
    dy = np.full((7, 7), 2.0, np.float32)
    dx[3, 3] = dy[3, 3] = np.nan
    dx[0, 0] = dy[0, 0] = np.nan   # corner: 3 neighbours < 6 needed
    fx, fy, filled = oracle.fill_holes(dx, dy)
    assert fx[3, 3] == 1 and fy[3, 3] == 2
    assert np.isnan(fx[0, 0])
    assert list(filled) == [3 + 3 * 7]


# --------------------------------------------------------------------------------------------
# pre-filter and quantizer
# --------------------------------------------------------------------------------------------

def test_highpass_vs_brute():
    rng = np.random.default_rng(8)
    img = (rng.random((40, 33)) * 255).astype(np.float32)
    mask = np.ones(img.shape, bool)
    out, vm = oracle.highpass_filtered(img, mask, 5)
    exp = img - _brute_window(img, 5, np.mean)
    np.testing.assert_allclose(out, exp, atol=1e-4)
    assert vm.all()
    mask[10:14, 5:9] = False
    out, vm = oracle.highpass_filtered(img, mask, 5)
    m = np.where(mask, img, np.nan)
    exp = np.where(mask, img - _brute_window(m, 5, np.mean), 0)
    np.testing.assert_allclose(out, exp, atol=1e-4)
    ero = _brute_window(np.where(mask, 1.0, 0.0), 5, np.min) > 0.5
    np.testing.assert_array_equal(vm, ero)
    assert np.all(out[~mask] == 0)   # invalid pixels filter to NaN, then to 0


def test_to_uint8():
    rng = np.random.default_rng(9)
    img = rng.normal(0, 10, (50, 40)).astype(np.float32)
    mask = np.ones(img.shape, bool)
    mask[0, :5] = False
    q = oracle.to_uint8(img, mask)
    v = img[mask].astype(np.float64)
    mean, sd = v.mean(), v.std(ddof=1)
    t = np.clip((img.astype(np.float64) - (mean - 3 * sd)) * (255 / (6 * sd)), 0, 255)
    exp = np.where(mask, np.rint(t), 0)
    # summation order may move a value across a rounding boundary only at ~1e-12
    assert np.sum(q != exp) == 0
    assert np.all(q[0, :5] == 0)
    flat = np.full((4, 4), 3.0, np.float32)
    assert np.all(oracle.to_uint8(flat, np.ones((4, 4), bool)) == 0x80)
    # round(UInt8, x) is ties-to-even: 2.5 -> 2, 3.5 -> 4
    assert np.rint(2.5) == 2 and np.rint(3.5) == 4


# --------------------------------------------------------------------------------------------
# whole passes
# --------------------------------------------------------------------------------------------

def _grid(n, spacing, margin):
    xs = np.arange(margin + 1, n - margin + 1, spacing, dtype=np.float64)
    X, Y = np.meshgrid(xs, xs)
    return X, Y


@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_track_exact_grid_recovery(dtype):
    # test/track.jl:15-52: integer shift recovered exactly at every point, sign flipped.
    n, shift = 200, (3, -2)
    ref, sec = shifted_pair(n, shift, seed=1, dtype=dtype)
    X, Y = _grid(n, 16, 45)
    ok = np.ones(ref.shape, bool)
    dx, dy, c, s = oracle.track(ref, sec, ok, X, Y, 25, 25, 32, 32, up=1)
    assert s.all()
    assert np.all(dx == -shift[0]) and np.all(dy == -shift[1])
    dx, dy, c, s = oracle.track(ref, sec, ok, X, Y, 25, 25, 32, 32, up=64)
    assert np.all(np.abs(dx + shift[0]) <= 3 / 64) and np.all(np.abs(dy + shift[1]) <= 3 / 64)


def test_track_threaded_is_bitwise_serial():
    ref, sec = shifted_pair(160, (1.25, -0.5), seed=2)
    X, Y = _grid(160, 12, 20)   # border forces padding
    ok = np.ones(ref.shape, bool)
    a = oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16, nthreads=1)
    b = oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16, nthreads=4)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_track_edges_masks_priors():
    ref, sec = shifted_pair(120, (0, 0), seed=3)
    ok = np.ones(ref.shape, bool)
    ok[:, :60] = False
    x = np.array([90.0, 30.0, -100.0, 60.0])
    y = np.array([60.0, 60.0, 60.0, 60.0])
    dx, dy, c, s = oracle.track(ref, sec, ok, x, y, 10, 10, 16, 16)
    assert s[0] and not s[1] and not s[2]   # masked chip and pure padding are skipped
    assert s[3]                               # chip straddling the mask edge is searched
    # a prior moves the chip; the reported displacement includes it
    ref, sec = shifted_pair(160, (5, 0), seed=4)
    dx, dy, c, s = oracle.track(ref, sec, np.ones(ref.shape, bool), np.array([80.0]),
                                np.array([80.0]), 6, 6, 32, 32, pdx=-5.0)
    assert dx[0] == -5 and dy[0] == 0
    # degenerate chip: searched, but no measurement
    flat = np.full((80, 80), 7, np.uint8)
    dx, dy, c, s = oracle.track(flat, flat, np.ones(flat.shape, bool), np.array([40.0]),
                                np.array([40.0]), 6, 6, 16, 16)
    assert s[0] and np.isnan(dx[0])


def test_multichip_resolves_shift():
    ref, sec = shifted_pair(256, (4, -3), seed=5, dtype=np.uint8)
    ok = np.ones(ref.shape, bool)
    X, Y = _grid(256, 8, 90)
    dx, dy, c, cs, interp = oracle.multichip(ref, sec, ok, X, Y, 25, 25,
                                             p=oracle.params(chip_size=32))
    meas = ~np.isnan(dx)
    assert meas.mean() > 0.9
    assert np.all(cs[meas] == 32)
    assert np.all(np.abs(dx[meas] + 4) <= 3 / 64) and np.all(np.abs(dy[meas] - 3) <= 3 / 64)


def test_quant_stats_is_the_sequential_sum():
    """orc_quant_stats (src/preprocess.jl:547-570) is a left-to-right Float64 sum, not a
    correctly rounded or pairwise one. Pinned by a case where the orders disagree: 2^53 followed
    by ones. Left to right every 2^53 + 1 rounds back to 2^53 (ties to even), so the ones vanish;
    the exact sum keeps them. The squared deviations are pinned the same way against a plain
    Python loop, whose float additions are sequential."""
    n = 1001
    v = np.ones(n)
    v[0] = 2.0 ** 53
    mask = np.ones(n, bool)
    cnt, mean, sd = oracle.quant_stats(v.reshape(-1, 1), mask.reshape(-1, 1))
    assert cnt == n
    assert mean == 2.0 ** 53 / n                      # the ones were absorbed
    assert mean != (2.0 ** 53 + (n - 1)) / n          # what an exact sum gives
    ss = 0.0
    for x in v:
        d = x - mean
        ss += d * d
    assert sd == np.sqrt(ss / (n - 1))
    # closed form on small integers: [1, 2, 3, 4] -> mean 2.5, sd sqrt(5/3); masked and
    # non-finite elements are skipped
    w = np.array([[1.0, np.nan], [2.0, 99.0], [3.0, np.inf], [4.0, 7.0]])
    m = np.array([[True, True], [True, False], [True, True], [True, False]])
    cnt, mean, sd = oracle.quant_stats(w, m)
    assert (cnt, mean, sd) == (4, 2.5, np.sqrt(5.0 / 3.0))
    assert oracle.quant_stats(w, np.zeros_like(m)) == (0, 0.0, 0.0)
    assert oracle.quant_stats(np.array([[5.0]]), np.array([[True]])) == (1, 5.0, 0.0)
