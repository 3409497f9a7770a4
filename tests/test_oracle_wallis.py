"""CPU pins of the oracle's Wallis pre-filter (src/preprocess.jl:158-159, 333-351, 416-438).

The oracle restates the reference's running-sum box means; these tests hold it against what the
definition fixes independently of that summation order: a brute-force window average over the
valid pixels in Float64 (population standard deviation about the measured mean), the min_std
floor, exclusion of masked and NaN pixels, flat windows, and the erosion / zeroing of _filtered.
A Bessel-corrected deviation, a missing floor, or masked pixels leaking into a window would each
fail one of them.
"""
import numpy as np

import oracle


def _brute(img, mask, w, min_std):
    """Direct definition: the box mean over the valid, non-NaN pixels of the clamped w x w window
    (left-biased margins for even extents), Float32 squared deviations about it, their box mean,
    then (x - mean) / max(sd, min_std)."""
    nr, nc = img.shape
    wx, wy = min(w, nc), min(w, nr)
    lx, ly = wx // 2, wy // 2
    rx, ry = wx - 1 - lx, wy - 1 - ly
    x = np.where(mask, img.astype(np.float32), np.nan).astype(np.float64)
    if mask.all():
        x = img.astype(np.float64)

    def boxmean(a):
        out = np.full(a.shape, np.nan)
        for i in range(nr):
            for j in range(nc):
                blk = a[max(i - ly, 0):min(i + ry, nr - 1) + 1, max(j - lx, 0):min(j + rx, nc - 1) + 1]
                v = blk[~np.isnan(blk)]
                if v.size:
                    out[i, j] = v.mean()
        return out

    mean = boxmean(x).astype(np.float32)
    ok = mask & np.isfinite(mean)
    d = np.where(ok, img.astype(np.float32) - mean, np.nan).astype(np.float32)
    sd = np.sqrt(np.maximum(boxmean((d * d).astype(np.float64)), 0.0))
    fl = np.float32(min_std)
    div = np.maximum(sd, fl)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = np.where(ok & np.isfinite(sd) & (div > 0),
                       (img.astype(np.float32) - mean) / div, np.nan)
    return out


def _eroded(mask, w):
    nr, nc = mask.shape
    lx = ly = w // 2
    out = np.zeros_like(mask)
    for i in range(nr):
        for j in range(nc):
            out[i, j] = mask[max(i - ly, 0):i + ly + 1, max(j - lx, 0):j + lx + 1].all()
    return out


def _scene(seed, shape=(23, 31)):
    rng = np.random.default_rng(seed)
    base = rng.standard_normal(shape).cumsum(0).cumsum(1)
    return (base * 3 + 100).astype(np.float32)


def _check(img, mask, w, min_std):
    got, vm = oracle.wallis_filtered(img, mask, w, min_std)
    ref = _brute(img, mask, w, min_std)
    evm = (mask if mask.all() else _eroded(mask, w)) & np.isfinite(ref)
    np.testing.assert_array_equal(vm, evm)
    assert np.all(got[~np.isfinite(ref)] == 0)        # non-finite outputs become 0
    fin = np.isfinite(ref)                             # (eroded pixels keep their values)
    np.testing.assert_allclose(got[fin], ref[fin], rtol=2e-4, atol=2e-4)
    return got, vm


def test_matches_brute_force_dense_and_masked():
    for seed, w, min_std in ((0, 5, 0.0), (1, 3, 0.0), (2, 7, 0.5), (3, 5, 2.0)):
        img = _scene(seed)
        _check(img, np.ones(img.shape, bool), w, min_std)
        m = np.random.default_rng(seed + 10).random(img.shape) > 0.1
        m[:, :3] = False
        _check(img, m, w, min_std)


def test_population_not_sample_deviation():
    """The squared deviations of each neighbour about its own local mean, averaged over the window
    with the window's count (no Bessel correction), on a 5 x 5 quadratic ramp, width 3."""
    img = (np.arange(25, dtype=np.float32).reshape(5, 5)) ** 2
    got, vm = oracle.wallis_filtered(img, np.ones((5, 5), bool), 3, 0.0)
    x = img.astype(np.float64)

    def win(a, i, j):
        return a[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2]

    mean = np.array([[win(x, i, j).mean() for j in range(5)] for i in range(5)])
    d2 = (x - mean) ** 2
    for i, j in ((0, 0), (2, 3), (4, 1), (1, 4)):
        w = win(d2, i, j)
        expect = (x[i, j] - mean[i, j]) / np.sqrt(w.mean())
        bessel = (x[i, j] - mean[i, j]) / np.sqrt(w.sum() / (w.size - 1))
        assert abs(got[i, j] - expect) < 2e-6 * max(1.0, abs(expect))
        assert abs(got[i, j] - bessel) > 1e-2 * abs(expect)


def test_min_std_floor():
    img = _scene(4)
    img = (img - img.mean()) * np.float32(1e-3) + np.float32(50)   # local sd ~ 1e-3
    ones = np.ones(img.shape, bool)
    g0, _ = oracle.wallis_filtered(img, ones, 5, 0.0)
    g1, v1 = oracle.wallis_filtered(img, ones, 5, 1.0)
    # with the floor the output is (x - mean) / 1: tiny; without it, O(1)
    assert np.abs(g1[v1]).max() < 0.05
    assert np.abs(g0).max() > 0.5


def test_flat_windows_are_invalid():
    img = np.full((12, 14), 7.0, np.float32)
    got, vm = oracle.wallis_filtered(img, np.ones(img.shape, bool), 5, 0.0)
    assert not vm.any() and np.all(got == 0)
    # with a floor a flat window is finite (x - mean = 0)
    got, vm = oracle.wallis_filtered(img, np.ones(img.shape, bool), 5, 0.25)
    assert vm.all() and np.all(got == 0)


def test_masked_and_nan_pixels_do_not_leak():
    img = _scene(5)
    m = np.ones(img.shape, bool)
    m[8:11, 12:15] = False
    a, va = oracle.wallis_filtered(img, m, 5, 0.0)
    img2 = img.copy()
    img2[~m] = 1e6          # whatever a masked pixel holds, no window sees it
    b, vb = oracle.wallis_filtered(img2, m, 5, 0.0)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    # a NaN pixel under an all-true mask is skipped by the means, and is itself invalid
    img3 = img.copy()
    img3[4, 6] = np.nan
    c, vc = _check(img3, np.ones(img.shape, bool), 5, 0.0)
    assert not vc[4, 6]


def test_affine_invariance():
    img = _scene(6)
    ones = np.ones(img.shape, bool)
    a, va = oracle.wallis_filtered(img, ones, 5, 0.0)
    b, vb = oracle.wallis_filtered((img * np.float32(4) - np.float32(37)).astype(np.float32),
                                   ones, 5, 0.0)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_allclose(a, b, atol=2e-4)
