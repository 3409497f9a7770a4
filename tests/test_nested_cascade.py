"""The nested-block settlement of the 64x subpixel cascade (CPU).

The refinement kernel first computes 16 x 16 blocks at 8x, 16x, 32x and 64x, each centred on the
previous block's argmax and placed so that every sample feeding it lies in that block (the 4x
level is whole). A source sample s of a pyrUp axis over n samples feeds exactly the outputs
[max(2s - 2, 0), min(2s + 2, 2n - 1)]; a sample is interior when all of them lie in the next
block. Every 64x sample outside the final block is then at most ub_margin of the largest
non-interior sample below, and when that bound is strictly below the final block's maximum the
block's argmax (peak_index rule) is the level's (src/peak.jl:32-61, 456-522). These tests check
the feeding rule against the oracle's pyrUp, the block placement, and the settlement against the
argmax of the whole 320 x 320 level on sharp, flat, edge, NaN and tied patches.
"""
import numpy as np
import pytest

import oracle

F32 = np.float32


def _fed(s, n):
    return max(2 * s - 2, 0), min(2 * s + 2, 2 * n - 1)


def _feed_range(a, b, n):
    return max(((a + 1) >> 1) - 1, 0), min((b >> 1) + 1, n - 1)


def _next_block(r0, i, n):
    r = 2 * i - 7
    lo = 0 if r0 == 0 else 2 * r0 + 1
    hi = 2 * n - 16 if r0 + 16 == n else 2 * r0 + 14
    r = min(max(r, lo), hi)
    return min(max(r, 0), 2 * n - 16)


def _argmax(a):
    v = a.ravel()
    ok = ~np.isnan(v)
    if not ok.any():
        return None, -np.inf
    m = v[ok].max()
    return int(np.flatnonzero(v == m)[0]), m


def _ub(m):
    return m if not m > -np.inf else F32(m) + F32(abs(m)) * F32(1e-6) + F32(1e-30)


def _ring(vals, r0, c0, n, R1, C1):
    m = -np.inf
    for r in range(vals.shape[0]):
        lo, hi = _fed(r0 + r, n)
        rin = lo >= R1 and hi <= R1 + 15
        for c in range(vals.shape[1]):
            lo2, hi2 = _fed(c0 + c, n)
            v = vals[r, c]
            if not (rin and lo2 >= C1 and hi2 <= C1 + 15) and not np.isnan(v) and v > m:
                m = v
    return m


def _settle(levels, F=6):
    """(settled, flat index) of the nested blocks from the 4x level."""
    z = levels[2]
    nb = z.shape[0]
    zi, _ = _argmax(z)
    if zi is None:
        return False, None
    ib, jb = divmod(zi, nb)
    R0 = min(max(2 * ib - 7, 0), 2 * nb - 16)
    C0 = min(max(2 * jb - 7, 0), 2 * nb - 16)
    mring = _ring(z, 0, 0, nb, R0, C0)
    src, sr0, sc0 = z, 0, 0
    for L in range(3, F + 1):
        n = 5 << L
        lo, hi = _feed_range(R0, R0 + 15, n // 2)
        lo2, hi2 = _feed_range(C0, C0 + 15, n // 2)
        assert lo >= sr0 and hi < sr0 + src.shape[0] and lo2 >= sc0 and hi2 < sc0 + src.shape[1]
        blk = levels[L][R0:R0 + 16, C0:C0 + 16]
        bi, bv = _argmax(blk)
        if bi is None:
            return False, None
        if L == F:
            if not _ub(mring) < bv:
                return False, None
            return True, (R0 + bi // 16) * n + C0 + bi % 16
        R1, C1 = _next_block(R0, R0 + bi // 16, n), _next_block(C0, C0 + bi % 16, n)
        mring = max(mring, _ring(blk, R0, C0, n, R1, C1))
        src, sr0, sc0, R0, C0 = blk, R0, C0, R1, C1
    return False, None


def test_fed_outputs_match_the_pyrup_taps():
    """Perturbing source s changes exactly the outputs [max(2s - 2, 0), min(2s + 2, 2n - 1)]."""
    rng = np.random.default_rng(3)
    for n in (5, 10, 20):
        base = rng.random((n, n)).astype(F32) + F32(1)
        out0 = oracle.pyrup(base)
        for s in range(n):
            p = base.copy()
            p[s, :] += F32(0.5)
            rows = np.flatnonzero(np.any(oracle.pyrup(p) != out0, axis=1))
            lo, hi = _fed(s, n)
            assert rows.min() >= lo and rows.max() <= hi and set(range(lo, hi + 1)) == set(rows)


def test_block_placement_keeps_feeders_inside():
    for n in (40, 80, 160):
        for r0 in range(n - 15):
            for i in range(r0, r0 + 16):
                r = _next_block(r0, i, n)
                lo, hi = _feed_range(r, r + 15, n)
                assert 0 <= r <= 2 * n - 16 and r0 <= lo and hi <= r0 + 15


def _patches():
    rng = np.random.default_rng(7)
    y, x = np.mgrid[-2:3, -2:3]
    for _ in range(40):
        cy, cx = rng.uniform(-0.6, 0.6, 2)
        k = rng.uniform(0.02, 0.6)
        yield (0.9 - k * (x - cx) ** 2 - k * rng.uniform(0.5, 2) * (y - cy) ** 2
               + rng.normal(0, 0.01, (5, 5))).astype(F32)
    yield np.full((5, 5), 0.5, F32)                                   # flat
    for cy, cx in ((-2, -2), (2, 2), (1.9, -2.3), (0, 2.4)):         # peaks at the edges
        yield (0.8 - 0.1 * (x - cx) ** 2 - 0.2 * (y - cy) ** 2).astype(F32)
    p = (0.8 - 0.1 * x ** 2 - 0.1 * y ** 2).astype(F32)
    p[0, 0] = np.nan
    yield p
    yield (-0.3 - 0.1 * x ** 2 - 0.05 * y ** 2).astype(F32)          # negative
    yield (0.8 - 1e-5 * (x ** 2 + y ** 2)).astype(F32)               # nearly flat
    yield (0.8 - 0.1 * (x - 0.5) ** 2 - 0.1 * y ** 2).astype(F32)    # symmetric tie


def test_settled_blocks_give_the_level_argmax():
    settled = 0
    for patch in _patches():
        levels = [patch]
        for _ in range(6):
            levels.append(oracle.pyrup(levels[-1]))
        ok, idx = _settle(levels)
        if ok:
            settled += 1
            full, _ = _argmax(levels[6])
            assert idx == full
    assert settled >= 40      # every sharp peak settles; flat ones fall back to the tiles
