"""The UInt8 quantiser's statistics on the device, bit for bit against the reference's
left-to-right Float64 sums (src/preprocess.jl:546-592, restated in the oracle).

The mean and the sum of squared deviations are sequential sums; the device reproduces them with
a parallel walk over runs of steps (exact-zone runs, runs inside one binade with ties settled by
parity, steps taken on their own, chunks summed in order). Each case below is built to exercise
one of those branches at scale: sums that never round, sums that round on every step, ties to
even that depend on the running sum's parity, frequent binade changes around zero, leading masked
regions, infinities and degenerate images."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


def _params(**kw):
    p = autorift.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _case(kind, shape, seed):
    rng = np.random.default_rng(seed)
    n = shape[0] * shape[1]
    if kind == "walk_f32":          # zero-mean random walk with fine bits: many binade changes
        v = (rng.standard_normal(n) * 30).astype(np.float32)
        v[rng.random(n) < 1e-3] = np.float32(3e-7)
    elif kind == "positive_f32":    # monotone sum that rounds on nearly every step
        v = (rng.random(n) * 255).astype(np.float32)
    elif kind == "ties_f32":        # steps exactly halfway between two multiples of the spacing
        v = rng.choice(np.array([2.0 ** -23, 3 * 2.0 ** -23, 5 * 2.0 ** -23, 2.0 ** -22, 1.0],
                                np.float32), n).astype(np.float32)
        v[0] = np.float32(2.0 ** 30)
    elif kind == "big_f64":         # wide dynamic range in Float64
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)
    elif kind == "u8":
        v = rng.integers(0, 256, n).astype(np.uint8)
    elif kind == "i16":
        v = rng.integers(-30000, 30000, n).astype(np.int16)
    elif kind == "masked_f32":      # a leading masked third, NaNs, one infinity masked out
        v = (rng.standard_normal(n) * 5 + 40).astype(np.float32)
        v[rng.random(n) < 0.01] = np.nan
    elif kind == "inf_f32":         # infinities inside the mask: skipped by the isfinite test
        v = (rng.random(n) * 100).astype(np.float32)
        v[n // 2] = np.inf
        v[n // 3] = -np.inf
    elif kind == "overflow_f64":    # the running sum overflows: mean and deviation infinite / NaN
        v = rng.standard_normal(n)
        v[rng.random(n) < 0.1] = 1e307
    elif kind == "flat_f32":
        v = np.full(n, 7.25, np.float32)
    elif kind == "one_valid_f32":
        v = (rng.random(n) * 100).astype(np.float32)
    else:
        raise ValueError(kind)
    img = v.reshape(shape, order="F")
    mask = None
    if kind == "masked_f32":
        mask = np.isfinite(img)
        mask.ravel(order="F")[: n // 3] = False
    if kind == "inf_f32":
        mask = np.ones(shape, bool)
    if kind == "one_valid_f32":
        mask = np.zeros(shape, bool)
        mask[shape[0] // 2, shape[1] // 2] = True
    return img, mask


KINDS = ["walk_f32", "positive_f32", "ties_f32", "big_f64", "u8", "i16", "masked_f32", "inf_f32",
         "overflow_f64", "flat_f32", "one_valid_f32"]


def _check(img, mask, img2, mask2, preprocess):
    p = _params(quantize=1, preprocess=preprocess)
    with autorift.Context(p) as ctx:
        ctx.set_images(img, img2, mask, mask2)
        stats = ctx.quantizer_stats()
        r, s, _, _ = ctx.get_pair()
    for im, m, st, got in ((img, mask, stats[0], r), (img2, mask2, stats[1], s)):
        m = np.isfinite(im.astype(np.float64)) if m is None else m
        if preprocess == 1:
            src, vm = oracle.highpass_filtered(im, m)
        else:
            src, vm = im, m
        n, mean, sd = oracle.quant_stats(src, vm)
        assert st[0] == n
        if n:
            assert np.float64(st[1]).view(np.uint64) == np.float64(mean).view(np.uint64) or (
                np.isnan(st[1]) and np.isnan(mean)), (st[1], mean)
            assert np.float64(st[2]).view(np.uint64) == np.float64(sd).view(np.uint64) or (
                np.isnan(st[2]) and np.isnan(sd)), (st[2], sd)
        if src.dtype in (np.float32, np.uint8, np.int16) or preprocess == 1:
            np.testing.assert_array_equal(got, oracle.to_uint8(src, vm))


@pytest.mark.parametrize("kind", KINDS)
def test_quantizer_stats_bit_exact(kind):
    shape = (517, 1031)   # 533 027 elements: 131 chunks of 4096, the last one partial
    img, mask = _case(kind, shape, 11)
    img2, mask2 = _case(kind, shape, 12)
    print(f"quantizer stats {kind} {shape}", flush=True)
    _check(img, mask, img2, mask2, preprocess=0)


@pytest.mark.parametrize("kind", ["walk_f32", "positive_f32", "masked_f32", "u8"])
def test_quantizer_stats_after_highpass(kind):
    """The default pipeline: the statistics of the high-pass output."""
    shape = (700, 650)
    img, mask = _case(kind, shape, 21)
    img2, mask2 = _case(kind, shape, 22)
    print(f"quantizer stats after high-pass {kind} {shape}", flush=True)
    _check(img, mask, img2, mask2, preprocess=1)


def test_quantizer_stats_large_walk():
    """16.8 M elements: many chunks and records, the walk's read-ahead across batches."""
    shape = (4096, 4100)
    img, _ = _case("walk_f32", shape, 31)
    img2, _ = _case("positive_f32", shape, 32)
    print(f"quantizer stats large {shape}", flush=True)
    _check(img, None, img2, None, preprocess=0)


def test_quantizer_stats_small_images():
    """Fewer elements than one chunk, one lane, and a single element."""
    for shape in ((1, 1), (3, 5), (64, 1), (70, 70)):
        img, _ = _case("walk_f32", shape, 41)
        img2, _ = _case("positive_f32", shape, 42)
        print(f"quantizer stats small {shape}", flush=True)
        _check(img, None, img2, None, preprocess=0)


def test_walk_takes_the_fast_path():
    """On a textured Float32 image through the default high-pass, the device sums nearly every
    element through runs: few walk items, at most one chunk added element by element (the
    running sum of squares climbs through a binade every few steps at first)."""
    from _util import synthetic_texture
    t = (synthetic_texture((1024, 1024), seed=3) * 255).astype(np.float32)
    t2 = (synthetic_texture((1024, 1024), seed=4) * 255).astype(np.float32)
    with autorift.Context(_params(quantize=1, preprocess=1)) as ctx:
        ctx.set_images(t, t2)
        stats = ctx.quantizer_stats(walk=True)
    print(f"walk statistics {stats}", flush=True)
    for st in stats:
        for items, in_order in st[4:]:
            assert items <= 256
            assert in_order <= 4096


def _left_to_right(v):
    s = 0.0
    for x in v.tolist():   # Python float additions: sequential, round to nearest even
        s += x
    return s


def _sequences(rng):
    out = {}
    v = np.ones(5000)
    v[0] = 2.0 ** 53
    out["absorbed_ones"] = v                              # 2^53 + 1 rounds back to 2^53
    out["squares"] = (rng.standard_normal(300_000) * 30) ** 2      # monotone, rounds every step
    t = np.full(70_000, 2.0 ** -23)
    t[rng.random(t.size) < 0.3] = 3 * 2.0 ** -23
    t[0] = 2.0 ** 30
    out["ties"] = t                                       # every step a tie on the 2^-22 grid
    k = rng.integers(-2 ** 21, 2 ** 21, 16384) * 2 + 1
    out["u8_energy"] = (k.astype(np.float64) ** 2) / 2.0 ** 28     # odd k: ties past 2^25
    out["tiny_then_large"] = np.concatenate([rng.random(3000) * 1e-300,
                                             rng.random(3000) * 1e300])
    return out


def test_sequential_sum_one_wave():
    """The one-wave replay the trackers use for chip energies (non-negative terms)."""
    rng = np.random.default_rng(5)
    with autorift.Context() as ctx:
        for name, v in _sequences(rng).items():
            print(f"one-wave sum {name} ({v.size})", flush=True)
            got, _, _ = ctx.sequential_sum(v, method=1)
            exp = _left_to_right(v)
            assert np.float64(got).tobytes() == np.float64(exp).tobytes(), (name, got, exp)


def test_sequential_sum_chunk_walk():
    """The chunk walk the quantiser uses, on signed and non-finite sequences as well."""
    rng = np.random.default_rng(6)
    seqs = _sequences(rng)
    seqs["walk_f32"] = (rng.standard_normal(2_000_003) * 30).astype(np.float32).astype(np.float64)
    w = rng.standard_normal(200_000) * 10.0 ** rng.integers(-30, 30, 200_000)
    seqs["signed_wide"] = w
    z = rng.standard_normal(50_000)
    z[::7] = -0.0
    z[1000:1100] = 5e-324
    seqs["zeros_denormals"] = z
    inf = rng.standard_normal(20_000)
    inf[15_000] = np.inf
    seqs["infinite"] = inf
    nan = rng.standard_normal(20_000)
    nan[100] = np.inf
    nan[200] = -np.inf
    seqs["nan"] = nan
    with autorift.Context() as ctx:
        for name, v in seqs.items():
            print(f"chunk-walk sum {name} ({v.size})", flush=True)
            got, runs, one_by_one = ctx.sequential_sum(v, method=0)
            exp = _left_to_right(v)
            assert np.float64(got).tobytes() == np.float64(exp).tobytes() or (
                np.isnan(got) and np.isnan(exp)), (name, got, exp)
            print(f"  runs {runs}, added one by one {one_by_one}", flush=True)


def test_sequential_sum_rejects_bad_input():
    with autorift.Context() as ctx:
        with pytest.raises(Exception):
            ctx.sequential_sum(np.array([1.0, -2.0]), method=1)
        with pytest.raises(Exception):
            ctx.sequential_sum(np.array([1.0]), method=2)
