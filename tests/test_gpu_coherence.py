"""GPU tests of the complex (SLC) path: Coherence tracking, Deramp and a measure per chip-size
level, against the oracle (itself pinned by tests/test_oracle_complex.py).

Coherence passes are bit-exact against the oracle: the kernel keeps the sequential Float64
order of the chip statistics, the summed-area tables and the direct complex numerator. Deramp
is held to a tolerance instead: the ramp sums are reduced in parallel and the device's sincos
and atan2 are not the C library's, so a deramped pixel may differ in its last Float32 bits.
"""
import numpy as np
import pytest

import oracle
from test_oracle_complex import speckle

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


def _pair(n, shift=(5, -3), seed=0xABCD, noise=0.05):
    z1 = speckle(n, seed=seed)
    z2 = np.roll(z1, shift, axis=(0, 1))
    rng = np.random.default_rng(seed + 1)
    z2 = (z2 + noise * (rng.standard_normal(z2.shape) + 1j * rng.standard_normal(z2.shape)))
    return z1, z2.astype(np.complex64)


def _ctx(**kw):
    kw.setdefault("quantize", "none")
    kw.setdefault("preprocess", "none")
    kw.setdefault("similarity", "coherence")
    return autorift.Context(autorift.params(**kw))


@pytest.mark.parametrize("cs,r,up", [(16, 6, 1), (32, 12, 16), (24, 9, 64), (32, 54, 16),
                                     (64, 64, 4)])
def test_track_matches_oracle(cs, r, up):
    """Interior, edge (virtual zero padding) and masked points; a flat chip is degenerate."""
    # (32, 54): window and chip within 1 KiB of the LDS limit -- the global-scratch instance
    # when the LDS one cannot be placed; (64, 64): a window far beyond LDS
    n = 180 if r < 50 else 360
    z1, z2 = _pair(n)
    z2[70:110, 70:110] = 3 - 2j          # flat: searched, no measurement
    m = np.ones(z1.shape, bool)
    m[:, 150:] = False
    xs = np.array([5.0, 30.0, 60.5, 90.0, 120.0, 155.0, 176.0])
    pts = autorift.gridpoints_xy(xs, xs, chip_size=cs, search_radius=r)
    print(f"coherence c{cs} r{r} up{up}", flush=True)
    with _ctx() as ctx:
        ctx.set_images(z1, z2, m, m)
        got = ctx.track(pts, upsampling=up, measure=autorift.AR_COHERENCE)
        pr, ps, pm, _ = ctx.get_pair()
    np.testing.assert_array_equal(pr, z1)
    exp = oracle.track(pr, ps, m, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                       pts.chip_size_y, up=up, measure=oracle.COHERENCE, nthreads=16)
    np.testing.assert_array_equal(got.searched, exp[3])
    for a, b in zip((got.dx, got.dy, got.correlation), exp[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    fin = np.isfinite(got.dx)
    assert fin.sum() >= 10
    if cs <= 40:   # the chip at (90, 90) lies inside the flat patch: degenerate
        centre = (pts.x == 90.0) & (pts.y == 90.0)
        assert got.searched[centre].all() and np.isnan(got.dx[centre]).all()


def test_deramp_matches_oracle():
    n = 200
    z1, z2 = _pair(n, noise=0.0)
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ramp = np.exp(1j * (0.9 * J - 0.35 * I))
    r1 = (z1 * ramp).astype(np.complex64)
    r2 = (z2 * ramp).astype(np.complex64)
    r1[10:20, 30:50] = np.nan            # excluded from the estimate, zeroed by NoQuantize
    print("deramp both / x / y", flush=True)
    for axis in ("both", "x", "y"):
        with _ctx(preprocess="deramp", deramp_axis=axis) as ctx:
            ctx.set_images(r1, r2)
            pr, ps, pm, _ = ctx.get_pair()
        m1 = np.isfinite(r1)
        e1 = oracle.deramp(np.where(m1, r1, 0), m1, axis=axis)
        e1[~m1] = 0
        np.testing.assert_array_equal(pm, m1)
        np.testing.assert_allclose(pr, e1, rtol=0, atol=2e-6 * np.abs(r1[m1]).max())
        assert not pr[~m1].any()
        if axis == "both":
            px, py = oracle.ramp_phase(pr, pm)
            assert abs(px) < 1e-4 and abs(py) < 1e-4


def test_autorift_recovers_a_circular_shift():
    """test/complex.jl:297-310: a circshift of (5, -3) reads dy = -5, dx = +3."""
    z1, z2 = _pair(384, noise=0.0)
    print("autorift coherence deramp", flush=True)
    out = autorift.autorift(z1, z2, similarity="coherence", preprocess="deramp",
                            quantize="none", chip_size=32, search_radius=20, chip_size_max=32)
    fin = np.isfinite(out.dx)
    assert fin.sum() > 0
    assert np.median(out.dy[fin]) == pytest.approx(-5, abs=0.5)
    assert np.median(out.dx[fin]) == pytest.approx(3, abs=0.5)


def test_multichip_matches_oracle():
    n = 260
    z1, z2 = _pair(n, noise=0.3)
    z2[:, 150:] = (np.random.default_rng(3).standard_normal((n, n - 150)) * 0.5).astype(
        np.complex64)   # decorrelated half: the coarser level takes over or nothing measures
    ok = np.ones(z1.shape, bool)
    xs = np.arange(40.0, 221.0, 10.0)
    grid = autorift.gridpoints_xy(xs, xs, chip_size=16, search_radius=12)
    print("multichip coherence", flush=True)
    with _ctx(chip_size=16, chip_size_max=32, upsampling=16) as ctx:
        ctx.set_images(z1, z2, ok, ok)
        g = ctx.run_multichip(grid)
    op = oracle.params(chip_size=16, chip_size_max=32, upsampling=16, measure=oracle.COHERENCE)
    e = oracle.multichip(z1, z2, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior, p=op, nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])
    np.testing.assert_array_equal(g.interpolated, e[4])
    assert g.nmeasured() > 0.3 * g.dx.size


def test_measure_per_level_on_a_real_pair():
    """similarity = ("ncc", "zncc"): NCC at the finest level, ZNCC above it, bit-exact."""
    from _util import shifted_pair
    ref, sec = shifted_pair(300, (-2.6, 1.9), seed=6, dtype=np.float32)
    ref[:, 150:] = ref[:, 150:] * 0.02 + 0.5
    sec[:, 150:] = sec[:, 150:] * 0.02 + 0.5
    ok = np.ones(ref.shape, bool)
    xs = np.arange(40.0, 261.0, 10.0)
    grid = autorift.gridpoints_xy(xs, xs, chip_size=16, search_radius=10)
    print("multichip ncc, zncc", flush=True)
    p = autorift.params(similarity=("ncc", "zncc"), quantize="none", preprocess="none",
                        chip_size=16, chip_size_max=64)
    with autorift.Context(p) as ctx:
        ctx.set_pair(ref, sec, ok, ok)
        g = ctx.run_multichip(grid)
    op = oracle.params(chip_size=16, chip_size_max=64, measures=(oracle.NCC, oracle.ZNCC))
    e = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                         grid.dx_prior, grid.dy_prior, p=op, nthreads=8)
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    np.testing.assert_array_equal(g.chip_size, e[3])
    # the finest level did use the tuple's first measure: ZNCC everywhere gives other values
    e_z = oracle.multichip(ref, sec, ok, grid.x, grid.y, grid.radius_x, grid.radius_y,
                           grid.dx_prior, grid.dy_prior,
                           p=oracle.params(chip_size=16, chip_size_max=64, measure=oracle.ZNCC),
                           nthreads=8)
    assert not np.array_equal(np.nan_to_num(e[2]), np.nan_to_num(e_z[2]))


def test_type_and_filter_errors():
    z1, z2 = _pair(96)
    real = np.abs(z1).astype(np.float32)
    pts = autorift.gridpoints_xy(np.array([48.0]), np.array([48.0]), chip_size=16,
                                 search_radius=6)
    with _ctx(preprocess="highpass") as ctx:
        with pytest.raises((autorift.AutoRIFTError, ValueError), match="amplitude filter"):
            ctx.set_images(z1, z2)
    with _ctx(quantize="uint8") as ctx:
        with pytest.raises((autorift.AutoRIFTError, ValueError), match="quantization"):
            ctx.set_images(z1, z2)
    with _ctx(preprocess="deramp") as ctx:
        with pytest.raises((autorift.AutoRIFTError, ValueError), match="complex input"):
            ctx.set_images(real, real)
    with _ctx() as ctx:
        ctx.set_images(real, real)
        with pytest.raises((autorift.AutoRIFTError, ValueError), match="needs complex input"):
            ctx.track(pts, upsampling=1, measure=autorift.AR_COHERENCE)
    with _ctx(similarity="zncc") as ctx:
        ctx.set_images(z1, z2)
        with pytest.raises((autorift.AutoRIFTError, ValueError), match="need real images"):
            ctx.track(pts, upsampling=1, measure=autorift.AR_ZNCC)


def test_served_complex_pair_equals_alone():
    """A complex pair through the serving loop gives the multi-chip result of a context."""
    z1, z2 = _pair(300, noise=0.2)
    p = autorift.params(similarity="coherence", preprocess="deramp", quantize="none",
                        chip_size=16, chip_size_max=32, search_radius=10)
    print("served coherence pair", flush=True)
    with autorift.Server(p, slots=2) as srv:
        t = srv.submit(z1, z2)
        served, _ = srv.result(t)
    alone = autorift.autorift(z1, z2, similarity="coherence", preprocess="deramp",
                              quantize="none", chip_size=16, chip_size_max=32, search_radius=10)
    for a, b in ((served.dx, alone.dx), (served.dy, alone.dy),
                 (served.correlation, alone.correlation)):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    assert np.isfinite(served.dx).mean() > 0.5
