"""Error bound of the u8 tracker's FP32 screening score (CPU).

The matrix-core path scores every shift in Float32 before evaluating the few shifts that can
still be the maximum with the reference's Float64 normalisation (src/correlate.jl:591-600); it
relies on |screening score - exact score| <= AR_SCREEN_EPS (4e-6, csrc/track.hip). This test
evaluates the kernel's screening arithmetic with numpy Float32 operations -- n num and n var as
exact integers, two conversions, the per-point constant 1 / (cnorm sqrt(n)) as v_rsq_f32 of
Float32(cnorm^2 n), v_rsq_f32 taken as the correctly rounded 1/sqrt moved by -1, 0 and +1 ulp
(its documented error is 1 ulp), two products -- and compares it with the oracle's exact surface (oracle.correlate, the reference's
own order) on windows built to stress it: random, nearly flat (variance tiny against the mean),
flat, saturated, and the largest chip area the path takes (128 x 128).
"""
import numpy as np
import pytest

import oracle

SCREEN_EPS = 4e-6   # AR_SCREEN_EPS in csrc/track.hip


def _shift_sums(search, chip):
    """Exact integer S1, S2 and the centred matrix-core accumulator per shift (i, j)."""
    ch, cw = chip.shape
    win = np.lib.stride_tricks.sliding_window_view(search.astype(np.int64), (ch, cw))
    t = chip.astype(np.int64)
    s1 = win.sum(axis=(2, 3))
    s2 = (win * win).sum(axis=(2, 3))
    acc = np.einsum("ijuv,uv->ij", win - 128, t - 128)
    return s1, s2, acc


def _screen(search, chip, measure, rsq_ulps):
    ch, cw = chip.shape
    n = ch * cw
    s1, s2, acc = _shift_sums(search, chip)
    sum_t = int(chip.astype(np.int64).sum())
    # exact integers (all below 2^53): n num and n var
    nn = n * acc + (128 * n - sum_t) * s1 + n * (128 * sum_t - 16384 * n)
    nv = n * s2 if measure == oracle.NCC else n * s2 - s1 * s1
    x = chip.astype(np.float64)
    sq = ((x * n - sum_t) ** 2).sum() / (n * n)          # cnorm^2
    # k = 1 / (cnorm sqrt(n)) = v_rsq_f32(Float32(cnorm^2 n)), off by the same ulps as below
    k = np.float32(1.0 / np.sqrt(np.float64(np.float32(sq * n))))
    nvf = np.maximum(nv.astype(np.float32), np.float32(1e-30))
    r = (1.0 / np.sqrt(nvf.astype(np.float64))).astype(np.float32)
    if rsq_ulps:
        to = np.float32(np.inf if rsq_ulps > 0 else 0.0)
        r = np.nextafter(r, to).astype(np.float32)
        k = np.nextafter(k, to).astype(np.float32)
    return (nn.astype(np.float32) * k) * r


def _check(search, chip, measure=oracle.ZNCC):
    ch, cw = chip.shape
    ry = (search.shape[0] - ch + 1) // 2
    rx = (search.shape[1] - cw + 1) // 2
    exact, ok = oracle.correlate(np.asfortranarray(search), np.asfortranarray(chip), rx, ry,
                                 measure=measure)
    assert ok
    worst = 0.0
    for ulps in (-1, 0, 1):
        approx = _screen(search, chip, measure, ulps)
        worst = max(worst, float(np.abs(approx.astype(np.float64) - exact.astype(np.float64)).max()))
    return worst


def _pair(rng, ch, cw, r, kind):
    h, w = ch + 2 * r - 1, cw + 2 * r - 1
    chip = rng.integers(0, 256, (ch, cw), dtype=np.uint8)
    if kind == "random":
        search = rng.integers(0, 256, (h, w), dtype=np.uint8)
    elif kind == "nearly_flat":
        search = np.full((h, w), 200, np.uint8)
        m = rng.random((h, w)) < 0.01
        search[m] = rng.integers(199, 202, m.sum(), dtype=np.uint8)
    elif kind == "flat_patches":
        search = rng.integers(0, 256, (h, w), dtype=np.uint8)
        search[: h // 2, : w // 2] = 77
    elif kind == "saturated":
        search = np.where(rng.random((h, w)) < 0.5, 0, 255).astype(np.uint8)
        chip = np.where(rng.random((ch, cw)) < 0.5, 0, 255).astype(np.uint8)
    else:
        raise ValueError(kind)
    return search, chip


@pytest.mark.parametrize("kind", ["random", "nearly_flat", "flat_patches", "saturated"])
@pytest.mark.parametrize("measure", [oracle.ZNCC, oracle.NCC])
def test_screen_bound_chip32(kind, measure):
    rng = np.random.default_rng(7)
    search, chip = _pair(rng, 32, 32, 25, kind)
    worst = _check(search, chip, measure)
    # the analysis gives < 4e-7 relative plus the exact value's own rounding
    assert worst < 1e-6 < SCREEN_EPS, worst


@pytest.mark.parametrize("shape", [(16, 16), (64, 32), (128, 128)])
def test_screen_bound_chip_sizes(shape):
    rng = np.random.default_rng(11)
    ch, cw = shape
    search, chip = _pair(rng, ch, cw, 6, "random")
    assert _check(search, chip) < 1e-6
    search, chip = _pair(rng, ch, cw, 6, "nearly_flat")
    assert _check(search, chip) < 1e-6


def test_screen_bound_catches_a_dropped_sqrt_n():
    # the check is sensitive: a screening constant without its sqrt(n) misses the bound
    rng = np.random.default_rng(3)
    search, chip = _pair(rng, 32, 32, 8, "random")
    exact, _ = oracle.correlate(np.asfortranarray(search), np.asfortranarray(chip), 8, 8)
    wrong = _screen(search, chip, oracle.ZNCC, 0) * np.float32(np.sqrt(32 * 32))
    assert np.abs(wrong - exact).max() > SCREEN_EPS
