"""GPU parity of the two box-sum paths of the u8 tracker: per-pass box images of W and W^2 and
per-point sliding sums. Both are exact integers, so each must match the oracle bit for bit;
AR_BOX_IMAGES forces one or the other (the library reads it on every pass)."""
import numpy as np
import pytest

import oracle
from _util import shifted_pair

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")


@pytest.fixture(scope="module")
def ctx():
    c = autorift.Context()
    yield c
    c.close()


@pytest.fixture(params=["always", "never"])
def box_mode(request, monkeypatch):
    monkeypatch.setenv("AR_BOX_IMAGES", request.param)
    return request.param


def _compare(ctx, ref, sec, pts, up, mode):
    ok = np.ones(ref.shape, bool)
    ctx.set_pair(ref, sec, ok, ok)
    print(f"box {mode} n={len(pts)} up={up} chip {pts.chip_size_y.flat[0]}x"
          f"{pts.chip_size_x.flat[0]}", flush=True)
    g = ctx.track(pts, upsampling=up)
    box_ms = ctx.last_timing()[0]
    assert (box_ms > 0) == (mode == "always")
    e = oracle.track(ref, sec, ok, pts.x, pts.y, pts.radius_x, pts.radius_y, pts.chip_size_x,
                     pts.chip_size_y, pts.dx_prior, pts.dy_prior, up=up, nthreads=8)
    np.testing.assert_array_equal(g.searched, e[3])
    for a, b in zip((g.dx, g.dy, g.correlation), e[:3]):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(a.view(np.uint32)[~np.isnan(a)],
                                      b.view(np.uint32)[~np.isnan(b)])
    assert ctx.last_kernel()[1] == 0
    return g


def test_box_grid_with_border_padding(ctx, box_mode):
    # windows cross the image edge, so the box images cover the zero-padded domain
    ref, sec = shifted_pair(300, (2.5, -1.25), seed=11, dtype=np.uint8)
    xs = np.arange(10, 292, 13, dtype=float)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    for up in (1, 64):
        g = _compare(ctx, ref, sec, pts, up, box_mode)
    assert g.searched.all()


def test_box_rectangular_chip_and_radii(ctx, box_mode):
    ref, sec = shifted_pair(260, (-3.0, 2.0), seed=12, dtype=np.uint8)
    xs = np.arange(70, 190, 9, dtype=float)
    ys = np.arange(50, 210, 11, dtype=float)
    pts = autorift.gridpoints_xy(xs, ys, chip_size=16, search_radius=25)
    pts = pts.rebuild(chip_size_x=np.full(pts.shape, 64, np.int64),
                      chip_size_y=np.full(pts.shape, 16, np.int64),
                      radius_x=np.full(pts.shape, 12, np.int64),
                      radius_y=np.full(pts.shape, 20, np.int64))
    _compare(ctx, ref, sec, pts, 16, box_mode)


def test_box_tall_chip(ctx, box_mode):
    # 256-row chip: eight rows per lane in the box-image kernel
    ref, sec = shifted_pair(420, (1.5, 2.25), seed=13, dtype=np.uint8)
    xs = np.arange(60, 360, 37, dtype=float)
    ys = np.arange(150, 270, 29, dtype=float)
    pts = autorift.gridpoints_xy(xs, ys, chip_size=64, search_radius=8)
    pts = pts.rebuild(chip_size_y=np.full(pts.shape, 256, np.int64))
    _compare(ctx, ref, sec, pts, 64, box_mode)


@pytest.mark.parametrize("ny", [200, 201, 202, 203])
def test_box_row_counts(ctx, box_mode, ny):
    # every residue of the column length mod 4: the window and chip columns start at each byte
    # alignment in the int8 copies the tracker fetches from (all windows inside the image)
    ref, sec = shifted_pair(210, (-1.75, 2.5), seed=14 + ny, dtype=np.uint8)
    ref = np.ascontiguousarray(ref[:ny, :197])
    sec = np.ascontiguousarray(sec[:ny, :197])
    xs = np.arange(45, 160, 9, dtype=float)
    ys = np.arange(45, ny - 44, 9, dtype=float)
    pts = autorift.gridpoints_xy(xs, ys, chip_size=32, search_radius=16)
    for up in (1, 64):
        g = _compare(ctx, ref, sec, pts, up, box_mode)
        assert g.searched.all()


def test_single_passes_rebuild_box_images(ctx, monkeypatch):
    """Only a multi-chip run shares box images between its passes; two plain passes over the same
    pair each build their own (and agree)."""
    monkeypatch.setenv("AR_BOX_IMAGES", "always")
    ref, sec = shifted_pair(300, (1.5, -2.0), seed=13, dtype=np.uint8)
    ok = np.ones(ref.shape, bool)
    ctx.set_pair(ref, sec, ok, ok)
    xs = np.arange(40, 260, 17, dtype=float)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=20)
    print("box images rebuilt per pass", flush=True)
    g1 = ctx.track(pts, upsampling=16)
    assert ctx.last_timing()[0] > 0
    g2 = ctx.track(pts, upsampling=16)
    assert ctx.last_timing()[0] > 0
    np.testing.assert_array_equal(g1.dx.view(np.uint32), g2.dx.view(np.uint32))
    np.testing.assert_array_equal(g1.dy.view(np.uint32), g2.dy.view(np.uint32))
