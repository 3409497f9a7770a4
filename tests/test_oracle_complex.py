"""The oracle's Coherence path and deramp (src/correlate.jl:632-735, src/integral.jl:172-200,
src/preprocess.jl:184-245), pinned the way the reference's own test/complex.jl pins them:
against what the mathematics fixes (gamma(T, T) = 1, a known shift, invariance to a global
phase, a known phase ramp recovered, deramp idempotent) and against an independent numerator
(numpy's complex FFT in double precision)."""
import numpy as np
import pytest

import oracle


def speckle(n, seed=0):
    """Circular Gaussian field smoothed by two 2x2 box passes with wrap-around (test/complex.jl's
    construction), complex64."""
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    for _ in range(2):
        z = (z + np.roll(z, 1, 0) + np.roll(z, 1, 1) + np.roll(np.roll(z, 1, 0), 1, 1)) / 4
    return z.astype(np.complex64)


@pytest.mark.parametrize("cs,r", [(16, 6), (24, 8), (32, 10)])
def test_coherence_of_a_chip_with_itself_is_one(cs, r):
    w = speckle(cs + 2 * r - 1, seed=cs)
    chip = w[r:r + cs, r:r + cs]          # zero displacement sits at (r + 1, r + 1), 1-based
    s, ok = oracle.correlate(w, chip, r, measure=oracle.COHERENCE)
    assert ok
    assert np.unravel_index(np.argmax(s), s.shape) == (r, r)
    assert s.max() == pytest.approx(1.0, abs=1e-5)
    tol = 2048 * np.finfo(np.float32).eps
    assert np.all((s >= -1 - tol) & (s <= 1 + tol))


@pytest.mark.parametrize("sy,sx", [(3, -2), (-4, 5), (0, 1)])
def test_coherence_finds_a_known_shift(sy, sx):
    cs, r = 16, 6
    w = speckle(cs + 2 * r - 1, seed=7)
    chip = w[r + sy:r + sy + cs, r + sx:r + sx + cs]
    s, ok = oracle.correlate(w, chip, r, measure=oracle.COHERENCE)
    assert np.unravel_index(np.argmax(s), s.shape) == (r + sy, r + sx)
    assert s.max() == pytest.approx(1.0, abs=1e-5)


def test_invariant_to_a_global_phase():
    cs, r = 16, 6
    w = speckle(cs + 2 * r - 1, seed=3)
    chip = speckle(cs, seed=4) + w[r:r + cs, r:r + cs]
    base, _ = oracle.correlate(w, chip, r, measure=oracle.COHERENCE)
    turned, _ = oracle.correlate(w, (chip * np.exp(1j * 1.234)).astype(np.complex64), r,
                                 measure=oracle.COHERENCE)
    np.testing.assert_allclose(base, turned, atol=1e-5)


def _numpy_surface(w, chip, r):
    """|sum conj(T') W| / (|T'| sqrt(sum |W|^2 - |sum W|^2 / n)) with the numerator from a
    zero-padded complex128 FFT cross-correlation and the window sums from direct box sums."""
    w = w.astype(np.complex128)
    c = chip.astype(np.complex128)
    ch, cw = c.shape
    t = c - c.mean()
    t = t.astype(np.complex64).astype(np.complex128)   # the chip is stored as ComplexF32
    cnorm = np.sqrt(np.sum(np.abs(c - c.mean()) ** 2))
    F = (w.shape[0] + ch, w.shape[1] + cw)
    num = np.fft.ifft2(np.fft.fft2(w, F) * np.conj(np.fft.fft2(t, F)))
    nr = nc = 2 * r
    out = np.zeros((nr, nc))
    n = ch * cw
    for i in range(nr):
        for j in range(nc):
            win = w[i:i + ch, j:j + cw]
            var = max(np.sum(np.abs(win) ** 2) - abs(win.sum()) ** 2 / n, 0.0)
            den = cnorm * np.sqrt(var)
            out[i, j] = abs(num[i, j]) / den if den > 0 else 0.0
    return out


@pytest.mark.parametrize("cs,r", [(16, 6), (20, 9)])
def test_direct_numerator_agrees_with_an_fft(cs, r):
    """test/complex.jl:162-186's agreement, against numpy's transform rather than FFTW's."""
    w = speckle(cs + 2 * r - 1, seed=11) * 30
    chip = (speckle(cs, seed=12) * 30 + w[r + 1:r + 1 + cs, r - 2:r - 2 + cs]).astype(
        np.complex64)
    s, ok = oracle.correlate(w.astype(np.complex64), chip, r, measure=oracle.COHERENCE)
    ref = _numpy_surface(w, chip, r)
    assert ok
    np.testing.assert_allclose(s, ref, rtol=2e-6, atol=2e-7)


def test_degenerate_chip_gives_zero_surface():
    cs, r = 16, 6
    w = speckle(cs + 2 * r - 1, seed=5)
    chip = np.full((cs, cs), 2 - 1j, np.complex64)
    s, ok = oracle.correlate(w, chip, r, measure=oracle.COHERENCE)
    assert not ok and not s.any()


def test_uncorrelated_speckle_is_low():
    cs, r = 32, 10
    w = speckle(cs + 2 * r - 1, seed=21)
    chip = speckle(cs, seed=22)
    s, _ = oracle.correlate(w, chip, r, measure=oracle.COHERENCE)
    # gamma of independent circular Gaussian fields ~ 1 / sqrt(N) per shift (N = 1024 samples,
    # smoothed: fewer independent ones); far below a match
    assert s.max() < 0.4 and np.all(np.isfinite(s))


def test_ramp_estimator_recovers_a_known_ramp():
    n = 64
    amp = np.abs(speckle(n, seed=8)) + 0.1
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for fx, fy in [(0.3, -0.2), (-1.1, 0.05), (2.0, 1.5)]:
        z = (amp * np.exp(1j * (fx * J + fy * I))).astype(np.complex64)
        px, py = oracle.ramp_phase(z)
        assert px == pytest.approx(fx, abs=1e-5)
        assert py == pytest.approx(fy, abs=1e-5)


def test_deramp_removes_the_ramp_and_is_idempotent():
    n = 64
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    z = (speckle(n, seed=9) * np.exp(1j * (0.7 * J - 0.4 * I))).astype(np.complex64)
    m = np.ones(z.shape, bool)
    d = oracle.deramp(z, m)
    px, py = oracle.ramp_phase(d, m)
    assert abs(px) < 1e-4 and abs(py) < 1e-4
    np.testing.assert_allclose(oracle.deramp(d, m), d, atol=1e-5)
    np.testing.assert_allclose(np.abs(d), np.abs(z), rtol=1e-5)
    # one axis: the other ramp stays (test/complex.jl:116-129)
    z2 = (speckle(n, seed=10) * np.exp(1j * 0.4 * I)).astype(np.complex64)
    dx_only = oracle.deramp(z2, m, axis="x")
    px2, py2 = oracle.ramp_phase(dx_only, m)
    assert abs(px2) < 1e-4
    assert py2 == pytest.approx(oracle.ramp_phase(z2, m)[1], abs=1e-7)
    # masked pixels are left out of the estimate but still deramped
    m2 = m.copy()
    m2[10:20, 10:20] = False
    z3 = z.copy()
    z3[10:20, 10:20] = 50 + 50j
    d3 = oracle.deramp(z3, m2)
    assert np.all(np.isfinite(d3)) and abs(d3[15, 15]) == pytest.approx(abs(z3[15, 15]), rel=1e-6)
    assert oracle.ramp_phase(z3, m2)[0] == pytest.approx(0.7, abs=0.02)


def test_coherence_needs_complex_and_real_measures_need_real():
    z = speckle(64, seed=1)
    ok = np.ones(z.shape, bool)
    with pytest.raises(ValueError):
        oracle.track(np.abs(z), np.abs(z), ok, 32.0, 32.0, 6, 6, 16, 16,
                     measure=oracle.COHERENCE)
    with pytest.raises(ValueError):
        oracle.track(z, z, ok, 32.0, 32.0, 6, 6, 16, 16, measure=oracle.ZNCC)
    with pytest.raises(ValueError):
        oracle.track(z, z, ok, 32.0, 32.0, 6, 6, 16, 16, measure=oracle.COHERENCE,
                     rotation=(-3.0, 0.0, 3.0))


def test_track_recovers_a_circular_shift():
    """test/complex.jl:297-310 at track level: a circshift of (5, -3) reads dy = -5, dx = +3."""
    n = 160
    z1 = speckle(n, seed=0xABCD)
    z2 = np.roll(z1, (5, -3), axis=(0, 1))
    ok = np.ones(z1.shape, bool)
    xs = np.arange(40.0, 121.0, 20.0)
    X, Y = np.meshgrid(xs, xs)
    for up in (1, 16):
        dx, dy, corr, srch = oracle.track(z1, z2, ok, X, Y, 12, 12, 32, 32, up=up,
                                          measure=oracle.COHERENCE)
        assert srch.all()
        np.testing.assert_allclose(dx, 3.0, atol=0.07)
        np.testing.assert_allclose(dy, -5.0, atol=0.07)
        if up == 1:   # the refined value is the pyrUp-smoothed peak, lower on sharp speckle
            assert np.all(corr > 0.999)


def test_multichip_with_a_measure_per_level():
    n = 256
    z1 = speckle(n, seed=0xABCD)
    z2 = np.roll(z1, (5, -3), axis=(0, 1))
    ok = np.ones(z1.shape, bool)
    xs = np.arange(50.0, 207.0, 16.0)
    X, Y = np.meshgrid(xs, xs)
    p = oracle.params(chip_size=16, chip_size_max=32, upsampling=16,
                      measures=(oracle.COHERENCE,))
    dx, dy, corr, cs, interp = oracle.multichip(z1, z2, ok, X, Y, 16, 16, p=p)
    fin = np.isfinite(dx)
    assert fin.mean() > 0.9
    assert np.median(dx[fin]) == pytest.approx(3.0, abs=0.5)
    assert np.median(dy[fin]) == pytest.approx(-5.0, abs=0.5)
