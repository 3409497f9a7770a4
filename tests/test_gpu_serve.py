"""GPU tests of the serving loop (ar_server): pairs of different sizes in flight together.

Every result must be the same alone and in the mix: the served pair runs the same kernels on
its own slot's buffers and stream, and the UInt8 path is exact, so "the same" is bit for bit.
"""
import numpy as np
import pytest

from bench import cfg1_pair

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")

SIZES = [512, 700, 384, 900, 640, 450, 1024, 560]


def _pairs():
    return [cfg1_pair(200 + k, n=n, shift=(1.5 + 0.25 * k, -0.75 * k)) for k, n in
            enumerate(SIZES)]


def _alone(p, pairs):
    out = []
    with autorift.Context(p) as ctx:
        for a, b in pairs:
            ctx.set_images(a, b)
            grid = autorift.api._build_grid(a.shape, p)
            out.append(ctx.run_multichip(grid))
    return out


def _same(r, e):
    for f in ("dx", "dy", "correlation"):
        g, x = getattr(r, f), getattr(e, f)
        assert g.shape == x.shape
        assert np.array_equal(g.view(np.uint32), x.view(np.uint32)), f
    np.testing.assert_array_equal(r.chip_size, e.chip_size)
    np.testing.assert_array_equal(r.interpolated, e.interpolated)


@pytest.mark.parametrize("prec", ["exact", "bf16"])
def test_results_same_alone_and_in_the_mix(prec):
    q = "uint8" if prec == "exact" else "none"
    p = autorift.params(quantize=q, precision=prec)
    pairs = _pairs()
    print(f"serve mix {prec}: alone", flush=True)
    alone = _alone(p, pairs)
    print(f"serve mix {prec}: served", flush=True)
    with autorift.Server(p, slots=4) as srv:
        tickets = [srv.submit(a, b) for a, b in pairs]
        got = [srv.result(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(pairs)
    assert st["peak_running"] >= 2          # pairs were on the device together
    assert len({t["slot"] for _, t in got}) >= 2
    for (r, tm), e in zip(got, alone):
        assert tm["latency_ms"] >= tm["service_ms"] > 0
        _same(r, e)
        assert r.nmeasured() > 0


def test_admission_by_device_memory():
    p = autorift.params()
    pairs = _pairs()[:4]
    alone = _alone(p, pairs)
    # footprint estimates of these pairs are 74-116 MB (the smallest two 155 MB together): a
    # 140 MB budget admits one at a time
    budget = 140 << 20
    print("serve admission", flush=True)
    with autorift.Server(p, slots=4, hbm_budget=budget) as srv:
        tickets = [srv.submit(a, b) for a, b in pairs]
        got = [srv.result(t) for t in tickets]
        st = srv.stats()
        assert st["peak_running"] == 1
        assert st["peak_bytes_held"] <= budget
        for (r, _), e in zip(got, alone):
            _same(r, e)
        # a pair whose footprint alone exceeds the budget is refused at submission
        big = np.zeros((4000, 4000), np.float32)
        with pytest.raises(autorift.ServerError):
            srv.submit(big, big)


def test_unknown_ticket_and_bad_input():
    with autorift.Server(autorift.params(), slots=1) as srv:
        with pytest.raises(autorift.ArgumentError):
            srv._pending[12345] = (1, 1)
            srv.result(12345)
        with pytest.raises(autorift.DimensionMismatch):
            srv.submit(np.zeros((64, 64), np.float32), np.zeros((64, 65), np.float32))


def test_served_pairs_with_masks():
    """Validity masks travel with a served pair: results equal the pair run alone with them."""
    p = autorift.params()
    pairs = _pairs()[:3]
    rng = np.random.default_rng(4)
    masks = [rng.random(a.shape) > 0.05 for a, _ in pairs]
    alone = []
    with autorift.Context(p) as ctx:
        for (a, b), m in zip(pairs, masks):
            ctx.set_images(a, b, m, m)
            alone.append(ctx.run_multichip(autorift.api._build_grid(a.shape, p)))
    print("serve masks", flush=True)
    with autorift.Server(p, slots=2) as srv:
        tickets = [srv.submit(a, b, m, m) for (a, b), m in zip(pairs, masks)]
        for t, e in zip(tickets, alone):
            r, _ = srv.result(t)
            _same(r, e)


def test_one_slot_gives_back_its_own_buffers():
    """One slot whose buffers from a larger pair exceed a small budget: it releases them and runs
    the next, smaller pair instead of waiting on itself (a pytest timeout would catch a hang)."""
    p = autorift.params()
    big, small = cfg1_pair(300, n=1400), cfg1_pair(301, n=384)
    with autorift.Context(p) as ctx:
        ctx.set_images(*small)
        exp = ctx.run_multichip(autorift.api._build_grid(small[0].shape, p))
    print("serve one slot, own buffers", flush=True)
    # a budget just above the large pair's footprint estimate (about 189 MiB)
    with autorift.Server(p, slots=1, hbm_budget=200 << 20) as srv:
        r0, _ = srv.result(srv.submit(*big))
        r1, _ = srv.result(srv.submit(*small))
        assert srv.stats()["completed"] == 2
    assert r0.nmeasured() > 0
    _same(r1, exp)
