"""Upper bound behind the refinement kernel's tile pruning (CPU).

The 64x subpixel cascade (src/peak.jl:196-522) skips a final-level tile only when a bound on every
sample it could hold is strictly below a sample already found. The bound: a pyrUp pass forms
0.75 b + 0.125 (a + c) or 0.5 (b + c) in FP32, which exceeds M = max(a, b, c) by at most
1.75 |M| 2^-24 (two passes per pyrUp step), so after the six passes from the 8x level to 64x a
sample exceeds the largest 8x sample feeding it by less than 6.3e-7 |M|; the kernel adds
1e-6 |M| + 1e-30 (ub_margin in csrc/refine.h). These tests run the oracle's pyrUp -- the same FP32
operations, pinned to OpenCV's pyrUp by the fixtures -- on inputs built to make the rounding
excess large, and check the per-step excess and the end-to-end margin.
"""
import numpy as np

import oracle

U = 2.0 ** -24
F32 = np.float32


def _feeders(k, n):
    """Source indices that can feed output k of a pyrUp axis of n source samples (a superset:
    the centre sample, its neighbours, and sample 1 for the reflected first output)."""
    c = k // 2
    s = {i for i in (c - 1, c, c + 1) if 0 <= i < n}
    if k == 0 and n > 1:
        s.add(1)
    return sorted(s)


def _stress_inputs(rng, n):
    """Values that maximise rounding in 0.75 b + 0.125 (a + c): mantissas near all ones, a mix of
    binades, signs, equal neighbours, tiny and subnormal magnitudes."""
    out = []
    near_one = F32(1) - np.arange(1, 9, dtype=np.float32) * F32(U)   # 1 - k ulp
    for _ in range(12):
        base = rng.choice(near_one, (n, n)) * F32(2.0 ** rng.integers(-20, 20))
        out.append(base.astype(F32))
    out.append((rng.random((n, n)) * 2 - 1).astype(F32))
    out.append(np.where(rng.random((n, n)) < 0.5, F32(1.0), np.nextafter(F32(1.0), F32(0))))
    out.append((rng.random((n, n)) * F32(1e-40)).astype(F32))        # subnormal
    out.append(np.full((n, n), F32(-3.0)))
    out.append((-rng.choice(near_one, (n, n))).astype(F32))
    return out


def test_pyrup_step_excess_is_bounded():
    rng = np.random.default_rng(11)
    worst = 0.0
    excess_seen = False
    for n in (5, 10, 20, 40):
        for src in _stress_inputs(rng, n):
            out = oracle.pyrup(src)
            s64 = src.astype(np.float64)
            for r in range(2 * n):
                fr = _feeders(r, n)
                for c in range(2 * n):
                    m = s64[np.ix_(fr, _feeders(c, n))].max()
                    v = float(out[r, c])
                    # two passes per step: at most 2 * 1.75 |M| u, plus subnormal rounding
                    assert v <= m + 3.5 * abs(m) * U + 4 * 2.0 ** -149, (n, r, c, v, m)
                    if v > m:
                        excess_seen = True
                        worst = max(worst, (v - m) / abs(m))
    # (on these inputs the rounding never actually lifts a sample above its feeders; the margin
    # is the proven worst case, not an observed one)
    assert worst <= 3.5 * U


def test_cascade_8x_to_64x_within_the_pruning_margin():
    """Every 64x sample is at most ub_margin(max of the 8x samples that can feed it)."""
    rng = np.random.default_rng(12)
    for trial in range(40):
        if trial % 2:
            patch = _stress_inputs(rng, 5)[trial % 17]
        else:
            patch = (rng.random((5, 5)) * 2 - 1).astype(F32)
        levels = [patch]
        for _ in range(6):
            levels.append(oracle.pyrup(levels[-1]))
        l8, l64 = levels[3].astype(np.float64), levels[6]
        n8 = l8.shape[0]
        # the 8x samples that can feed a 16-sample run of the 64x axis: chase the feeders back
        # through the 32x (n8 4 samples), 16x (n8 2) and 8x (n8) axes
        def feed(k0):
            a = set(range(k0, k0 + 16))
            for n in (n8 * 4, n8 * 2, n8):
                a = {f for j in a for f in _feeders(j, n)}
            return sorted(a)
        for t0 in range(0, 320, 16):
            rows = feed(t0)
            for u0 in range(0, 320, 16):
                m = F32(l8[np.ix_(rows, feed(u0))].max())
                bound = m + F32(abs(m)) * F32(1e-6) + F32(1e-30)   # ub_margin, in FP32
                tile = l64[t0:t0 + 16, u0:u0 + 16]
                assert tile.max() <= bound, (trial, t0, u0, tile.max(), bound)


def test_cascade_4x_to_64x_within_the_margin():
    """The nested-block path bounds 64x samples from the 4x level: eight passes, at most
    (1 + 1.75 u)^8 - 1 < 8.4e-7 |M| above the largest 4x feeder, inside ub_margin's 1e-6."""
    rng = np.random.default_rng(13)
    for trial in range(24):
        patch = (_stress_inputs(rng, 5)[trial % 17] if trial % 2 else
                 (rng.random((5, 5)) * 2 - 1).astype(F32))
        levels = [patch]
        for _ in range(6):
            levels.append(oracle.pyrup(levels[-1]))
        l4, l64 = levels[2].astype(np.float64), levels[6]
        n4 = l4.shape[0]

        def feed(k0):
            a = set(range(k0, k0 + 16))
            for n in (n4 * 8, n4 * 4, n4 * 2, n4):
                a = {f for j in a for f in _feeders(j, n)}
            return sorted(a)
        for t0 in range(0, 320, 32):
            rows = feed(t0)
            for u0 in range(0, 320, 32):
                m = F32(l4[np.ix_(rows, feed(u0))].max())
                bound = m + F32(abs(m)) * F32(1e-6) + F32(1e-30)
                assert l64[t0:t0 + 16, u0:u0 + 16].max() <= bound, (trial, t0, u0)
