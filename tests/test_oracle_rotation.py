"""The oracle's RotationSearch (src/track.jl:316-409, src/types.jl:426-577), pinned by closed forms
and by the reference's own behavioural test (test/firstguess.jl:244-290)."""
import numpy as np
import pytest

import oracle
from _util import chip_and_window, synthetic_texture


def test_zero_angle_is_identity_inside():
    """At 0 degrees every destination pixel samples itself; the last row and column have no
    bilinear neighbour inside the chip and are zero (the reference returns the chip unrotated at
    exactly 0, which the track loop does before rotating)."""
    rng = np.random.default_rng(0)
    chip = rng.random((12, 10)).astype(np.float32)
    out = oracle.rotate_chip(chip, 0.0)
    np.testing.assert_array_equal(out[:-1, :-1], chip[:-1, :-1])
    assert not out[-1, :].any() and not out[:, -1].any()


def test_quarter_turn_maps_pixels():
    """At 90 degrees destination (i, j) samples source (j, n + 1 - i) (1-based): the inverse
    mapping sy = yc + s dx, sx = xc - s dy with c = 0, s = 1 about the centre (n + 1) / 2."""
    n = 16
    rng = np.random.default_rng(1)
    chip = rng.random((n, n))
    out = oracle.rotate_chip(chip, 90.0)
    for i in range(2, n):            # 1-based interior rows / columns
        for j in range(2, n):
            assert out[i - 1, j - 1] == pytest.approx(chip[j - 1, n - i], abs=1e-6)


def test_bilinear_is_exact_on_a_linear_field():
    """Bilinear interpolation reproduces a linear field, so a rotated ramp is the ramp at the
    rotated coordinates; the centre is fixed and a 30-degree turn moves the gradient by 30
    degrees."""
    n = 24
    I, J = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    a, b = 0.7, -1.3
    chip = a * I + b * J + 5.0
    deg = 30.0
    out = oracle.rotate_chip(chip, deg).astype(np.float64)
    th = np.deg2rad(deg)
    yc = xc = (n + 1) / 2
    # a linear field rotated about the centre: value(i, j) = a sy + b sx + 5 where (sy, sx) is
    # the destination pixel turned by -deg in (row, column) coordinates
    sy = yc + np.cos(th) * (I - yc) + np.sin(th) * (J - xc)
    sx = xc - np.sin(th) * (I - yc) + np.cos(th) * (J - xc)
    inside = (np.floor(sy) >= 1) & (np.floor(sx) >= 1) & (np.floor(sy) + 1 <= n) & \
             (np.floor(sx) + 1 <= n)
    np.testing.assert_allclose(out[inside], (a * sy + b * sx + 5.0)[inside], rtol=0, atol=2e-5)
    assert not out[~inside].any()
    # the gradient of the rotated field, from two interior pixels along a row and a column
    c = n // 2
    gx = out[c, c + 1] - out[c, c]
    gy = out[c + 1, c] - out[c, c]
    assert gy == pytest.approx(a * np.cos(th) - b * np.sin(th), abs=1e-4)
    assert gx == pytest.approx(a * np.sin(th) + b * np.cos(th), abs=1e-4)


def _rotate_image(img, deg):
    """The scene rotated about its centre (test/utils.jl:208-224's arrangement, via the
    oracle's chip rotation applied to the whole image)."""
    return oracle.rotate_chip(img, deg)


def test_counter_rotation_recovers_correlation():
    """A chip cut from a secondary rotated by +8 degrees, against the unrotated reference: turning
    it back (-8) beats turning it further (+8) at every chip size, and beats not turning it at
    64 and 128 (test/firstguess.jl:244-290)."""
    n = 512
    a = synthetic_texture((n, n), seed=3).astype(np.float32)
    motion = 8.0
    b = _rotate_image(a, motion)
    ci = cj = 257
    R = 4

    def peak(cs, ang):
        chip, win = chip_and_window(b, (ci, cj), cs, R, window_from=a)
        c = np.asfortranarray(chip, dtype=np.float32)
        if ang != 0.0:
            c = oracle.rotate_chip(c, ang)
        s, ok = oracle.correlate(np.asfortranarray(win, dtype=np.float32), c, R, R)
        assert ok
        return float(s.max())

    for cs in (32, 64, 128):
        back, unrot, wrong = peak(cs, -motion), peak(cs, 0.0), peak(cs, motion)
        assert back > wrong, (cs, back, wrong)
        if cs >= 64:
            assert back > unrot, (cs, back, unrot)


def test_rotation_search_in_track():
    """angles (0,) is the unrotated pass bit for bit; on an unrotated scene (-3, 0, 3) keeps the
    unrotated surface wherever it has the highest peak, and every result is one of the three
    single-angle passes' results."""
    n = 160
    img = (synthetic_texture((n + 4, n + 4), seed=7) * 255).astype(np.float32)
    ref = img[2:n + 2, 2:n + 2].copy()
    sec = img[:n, 1:n + 1].copy()
    ok = np.ones(ref.shape, bool)
    xs = np.arange(40.0, 121.0, 20.0)
    X, Y = np.meshgrid(xs, xs)
    base = oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16)
    zero = oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16, rotation=(0.0,))
    for u, v in zip(base, zero):
        np.testing.assert_array_equal(u, v)
    rot = oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16, rotation=(-3.0, 0.0, 3.0))
    singles = [oracle.track(ref, sec, ok, X, Y, 10, 10, 32, 32, up=16, rotation=(a,))
               for a in (-3.0, 0.0, 3.0)]
    for k in np.ndindex(X.shape):
        cands = [(s[0][k], s[1][k], s[2][k]) for s in singles]
        assert (rot[0][k], rot[1][k], rot[2][k]) in cands
    # the scene is not rotated: the unrotated chip correlates best at almost every point
    same = np.mean((rot[0] == base[0]) & (rot[1] == base[1]))
    assert same >= 0.9
