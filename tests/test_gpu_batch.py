"""GPU tests of the batched pass (ar_batch): points of many pairs, uneven chip sizes and radii,
in one tracking pass.

Power-of-two chip areas must equal a pass over each pair alone, bit for bit (the same exact integer
arithmetic; the pass's shared padding only adds invalid zeros). Other chip areas do too when the
batch is exact (the same direct Float64 tracker); by default they take the integer tracker with
the unrounded mean-removed chip and are held to a tolerance.
"""
import numpy as np
import pytest

# PyTorch (device buffers for the device-pointer entry points) brings its own HIP runtime: it
# must be loaded before the library's first HIP call in a process that uses both.
torch = pytest.importorskip("torch")

from bench import cfg1_pair  # noqa: E402

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")

SIZES = [(300, 420), (512, 512), (640, 380), (450, 700)]


def _scene(seed=0, chips=(16, 32, 64), n_per=60):
    rng = np.random.default_rng(seed)
    pairs, pts, pair = [], [], []
    for k, (ny, nx) in enumerate(SIZES):
        a, b = cfg1_pair(300 + k, n=max(ny, nx), shift=(rng.uniform(-6, 6), rng.uniform(-6, 6)))
        pairs.append((a[:ny, :nx].copy(), b[:ny, :nx].copy()))
        x = rng.uniform(-5, nx + 5, n_per)       # some points near or past the edges
        y = rng.uniform(-5, ny + 5, n_per)
        cs = rng.choice(chips, n_per)
        r = rng.integers(8, 41, n_per)
        pts.append((x, y, cs, r))
        pair.append(np.full(n_per, k, np.int32))
    x = np.concatenate([p[0] for p in pts])
    y = np.concatenate([p[1] for p in pts])
    cs = np.concatenate([p[2] for p in pts]).astype(np.int64)
    r = np.concatenate([p[3] for p in pts]).astype(np.int64)
    ps = autorift.pointset(x, y, search_radius_x=r, search_radius_y=r, chip_size_x=cs,
                           chip_size_y=cs)
    return pairs, ps, np.concatenate(pair)


def _alone(pairs, ps, pair, up):
    """Each pair's points through a single-pair pass."""
    n = len(ps)
    out = [np.full(n, np.nan, np.float32) for _ in range(3)] + [np.zeros(n, bool)]
    with autorift.Context(autorift.params()) as ctx:
        for k, (a, b) in enumerate(pairs):
            sel = np.nonzero(pair == k)[0]
            ctx.set_images(a, b)
            r = ctx.track(ps[sel], upsampling=up)
            for o, v in zip(out, (r.dx, r.dy, r.correlation, r.searched)):
                o[sel] = v
    return out


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("up", [1, 64])
def test_batch_equals_each_pair_alone(up):
    pairs, ps, pair = _scene(1)
    print(f"batch vs alone up{up}: {len(ps)} points over {len(pairs)} pairs", flush=True)
    exp = _alone(pairs, ps, pair, up)
    with autorift.Batch(autorift.params()) as bt:
        for a, b in pairs:
            bt.add_pair(a, b)
        got = bt.track(ps, pair, upsampling=up)
        info = bt.info()
    assert info["pairs"] == len(pairs) and info["launches"] >= 1
    np.testing.assert_array_equal(got.searched, exp[3])
    for g, e in zip((got.dx, got.dy, got.correlation), exp[:3]):
        np.testing.assert_array_equal(_bits(g), _bits(e))
    assert np.count_nonzero(np.isfinite(got.dx)) > len(ps) // 2


def test_batch_splits_and_graph_replay():
    pairs, ps, pair = _scene(2)
    dev = torch.device("cuda:0")
    with autorift.Batch(autorift.params()) as bt:
        for a, b in pairs:
            bt.add_pair(a, b)
        ref = bt.track(ps, pair, upsampling=64)
        n = len(ps)
        flds = [torch.from_numpy(np.ascontiguousarray(np.ravel(getattr(ps, f), order="F"))).to(dev)
                for f in ("x", "y", "radius_x", "radius_y", "dx_prior", "dy_prior",
                          "chip_size_x", "chip_size_y")]
        pr = torch.from_numpy(pair.astype(np.int32)).to(dev)
        out = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3)]
        out.append(torch.empty(n, dtype=torch.uint8, device=dev))
        fp = [t.data_ptr() for t in flds]
        op = [t.data_ptr() for t in out]

        def check(tag):
            torch.cuda.synchronize()
            for t, e in zip(out[:3], (ref.dx, ref.dy, ref.correlation)):
                np.testing.assert_array_equal(_bits(t.cpu().numpy()), _bits(e), err_msg=tag)
            np.testing.assert_array_equal(out[3].cpu().numpy().astype(bool), ref.searched)

        for split, launches in ((0, 1), (16, 2), (24, 2), (32, 2), ((16, 28), 3), ((20, 32), 3)):
            print(f"batch split {split}", flush=True)
            for t in out:
                t.zero_()
            bt.track_device(fp, pr.data_ptr(), n, 64, 0, op, split=split)
            check(f"split {split}")
            assert bt.info()["launches"] == launches
        print("batch graph capture and replays", flush=True)
        for rep in range(3):
            for t in out:
                t.zero_()
            torch.cuda.synchronize()
            bt.track_device(fp, pr.data_ptr(), n, 64, 0, op, graph=True, split=24)
            assert bt.info()["replayed"] == (rep > 0)
            check(f"replay {rep}")


@pytest.mark.parametrize("up", [1, 64])
def test_batch_non_power_of_two_chips(up):
    """Chip 48 (area 2304) mixed with power-of-two chips in an exact batch: the direct Float64
    tracker inside the batched pass, bit-equal to each pair alone, also when the pass is captured
    and replayed."""
    pairs, ps, pair = _scene(3, chips=(16, 48, 32, 24), n_per=50)
    print(f"batch chip 48 exact up{up}", flush=True)
    exp = _alone(pairs, ps, pair, up)
    with autorift.Batch(autorift.params(), exact=True) as bt:
        for a, b in pairs:
            bt.add_pair(a, b)
        got = bt.track(ps, pair, upsampling=up)
        n = len(ps)
        dev = torch.device("cuda:0")
        flds = [torch.from_numpy(np.ascontiguousarray(np.ravel(getattr(ps, f), order="F"))).to(dev)
                for f in ("x", "y", "radius_x", "radius_y", "dx_prior", "dy_prior",
                          "chip_size_x", "chip_size_y")]
        pr = torch.from_numpy(pair.astype(np.int32)).to(dev)
        out = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3)]
        out.append(torch.empty(n, dtype=torch.uint8, device=dev))
        fp = [t.data_ptr() for t in flds]
        op = [t.data_ptr() for t in out]
        for rep in range(3):
            for t in out:
                t.zero_()
            torch.cuda.synchronize()
            bt.track_device(fp, pr.data_ptr(), n, up, 0, op, graph=True, split=24)
            assert bt.info()["replayed"] == (rep > 0)
            torch.cuda.synchronize()
            for t, e in zip(out[:3], exp[:3]):
                np.testing.assert_array_equal(_bits(t.cpu().numpy()), _bits(e))
    np.testing.assert_array_equal(got.searched, exp[3])
    for g, e in zip((got.dx, got.dy, got.correlation), exp[:3]):
        np.testing.assert_array_equal(_bits(g), _bits(e))
    assert np.isfinite(got.dx[ps.chip_size_x == 48]).mean() > 0.5


def test_batch_non_power_of_two_chips_default():
    """The default batch: chips 24 and 48 on the integer tracker (with the power-of-two chips),
    integer peaks equal to the exact pass, subpixel offsets within 0.05 px; switching the batch
    to exact and back gives each mode's own results."""
    pairs, ps, pair = _scene(3, chips=(16, 48, 32, 24), n_per=50)
    print("batch chip 48 default", flush=True)
    exp1 = _alone(pairs, ps, pair, 1)
    exp64 = _alone(pairs, ps, pair, 64)
    with autorift.Batch(autorift.params()) as bt:
        for a, b in pairs:
            bt.add_pair(a, b)
        g1 = bt.track(ps, pair, upsampling=1)
        g64 = bt.track(ps, pair, upsampling=64)
        bt.set_exact(True)
        x64 = bt.track(ps, pair, upsampling=64)
        bt.set_exact(False)
        d64 = bt.track(ps, pair, upsampling=64)
    for f, e in zip(("dx", "dy", "correlation"), exp64[:3]):
        np.testing.assert_array_equal(_bits(getattr(x64, f)), _bits(e))
        np.testing.assert_array_equal(_bits(getattr(d64, f)), _bits(getattr(g64, f)))
    np.testing.assert_array_equal(g64.searched, exp64[3])
    area = ps.chip_size_x * ps.chip_size_y
    p2 = (area & (area - 1)) == 0          # chips 16 and 32; 24 and 48 are not powers of two
    for f, e in zip(("dx", "dy", "correlation"), exp64[:3]):
        np.testing.assert_array_equal(_bits(getattr(g64, f)[p2]), _bits(e[p2]))
    m = ~p2 & np.isfinite(exp1[0])
    assert np.array_equal(np.isfinite(g1.dx[~p2]), np.isfinite(exp1[0][~p2]))
    assert ((g1.dx[m] == exp1[0][m]) & (g1.dy[m] == exp1[1][m])).all()
    ok = ~p2 & np.isfinite(exp64[0])
    off = np.hypot(g64.dx[ok] - exp64[0][ok], g64.dy[ok] - exp64[1][ok])
    print(f"chips 24, 48 default: {ok.sum()} points, max subpixel offset {off.max():.2e} px",
          flush=True)
    assert off.max() <= 0.05


def test_batch_errors():
    pairs, ps, pair = _scene(4, n_per=5)
    with autorift.Batch(autorift.params()) as bt:
        with pytest.raises(autorift.ArgumentError):
            bt.track(ps, pair)               # no pair yet
        bt.add_pair(*pairs[0])
        with pytest.raises(autorift.ArgumentError):
            bt.track(ps, pair)               # indices 1..3 are not in the batch
    with pytest.raises(autorift.AutoRIFTError):
        autorift.Batch(autorift.params(quantize="none"))


def test_batch_with_masks_and_gaps():
    """Pairs with invalid pixels (NaN gaps and explicit masks): each point's result equals a pass
    over its own pair alone, including the validity test of its chip (src/track.jl:416-421)."""
    pairs, ps, pair = _scene(5, n_per=50)
    masks = []
    rng = np.random.default_rng(9)
    for k, (a, b) in enumerate(pairs):
        a = a.copy()
        b = b.copy()
        ny, nx = a.shape
        if k % 2 == 0:
            a[ny // 3:ny // 3 + 40, :nx // 2] = np.nan          # a gap: non-finite pixels
            masks.append(None)
        else:
            m = rng.random((ny, nx)) > 0.02
            m[:, : nx // 4] = False                          # a masked-off strip
            masks.append(m)
        pairs[k] = (a, b)
    print("batch masks", flush=True)
    exp = [np.full(len(ps), np.nan, np.float32) for _ in range(3)] + [np.zeros(len(ps), bool)]
    with autorift.Context(autorift.params()) as ctx:
        for k, (a, b) in enumerate(pairs):
            sel = np.nonzero(pair == k)[0]
            ctx.set_images(a, b, masks[k], masks[k])
            r = ctx.track(ps[sel], upsampling=64)
            for o, v in zip(exp, (r.dx, r.dy, r.correlation, r.searched)):
                o[sel] = v
    with autorift.Batch(autorift.params()) as bt:
        for k, (a, b) in enumerate(pairs):
            bt.add_pair(a, b, masks[k], masks[k])
        got = bt.track(ps, pair, upsampling=64)
    np.testing.assert_array_equal(got.searched, exp[3])
    for g, e in zip((got.dx, got.dy, got.correlation), exp[:3]):
        np.testing.assert_array_equal(_bits(g), _bits(e))
    assert (~got.searched).any() and got.searched.any()
