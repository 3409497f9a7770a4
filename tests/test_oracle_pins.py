"""Known-answer pins for the oracle's grid-level steps: Gardner stages, hole filling, the coarse
gate and the merge.

No fixture of the reference covers these, and they are all simple arithmetic in grid
coordinates. So each test builds a grid by hand, small enough for the expected result to be
worked out with pen and paper, and writes that result into the test. The case is chosen so
that a plausible slip fails it: `<=` for `<`, a lower-middle median, a missing quantisation
floor, in-place instead of two-phase filling, swapped cell margins, and the like.
"""
import numpy as np

import oracle

F32 = np.float32


def _gardner(dx, dy, r, up=64, **kw):
    dx = np.asarray(dx, F32)
    dy = np.asarray(dy, F32)
    rr = np.full(dx.shape, r, np.int64)
    valid = ~np.isnan(dx)
    return oracle.reject_gardner(dx, dy, rr, rr, valid, up, oracle.gardner(**kw))


# --------------------------------------------------------------------------------------------
# Gardner stage 2: tolerance max(mad_scale * MAD, 2 / (up * r))  -- src/outliers.jl:138-153
# --------------------------------------------------------------------------------------------

def test_stage2_quantisation_floor():
    # A flat field (MAD = 0 everywhere) with one point off by d px. The neighbourhood median is
    # 0 and the MAD is 0, so the tolerance is the floor 2 / (up r) in normalised units, which is
    # exactly 2 / up px. A point exactly on the floor is kept (<=); one float32 ulp above it is
    # rejected. r and up are powers of two, so d / r is exact.
    for r in (16, 32):
        for up in (16, 64):
            edge = F32(2.0 / up)
            for d, kept in ((edge, True), (np.nextafter(edge, F32(1)), False),
                            (np.nextafter(edge, F32(0)), True)):
                for axis in (0, 1):
                    dx = np.zeros((7, 7), F32)
                    dy = np.zeros((7, 7), F32)
                    (dx, dy)[axis][3, 3] = d
                    keep = _gardner(dx, dy, r, up=up, iterations=2)
                    exp = np.ones((7, 7), bool)
                    exp[3, 3] = kept
                    np.testing.assert_array_equal(keep, exp, err_msg=f"r={r} up={up} d={d}")


def test_stage2_mad_scale():
    # 5x5 grid: the centre's window is the whole grid. Values (normalised, r = 32):
    # 8 at -a, 8 at 0, 8 at +a and the centre c > a. Sorted, the 13th of 25 is 0, so the median
    # is 0; the deviations are 8 zeros, 16 times a and c, so the MAD is a. The centre is kept
    # exactly when |c| <= mad_scale * a (the floor 2^-10 is below that).
    a = F32(2.0 ** -6)
    base = np.array([-a] * 8 + [F32(0)] * 8 + [a] * 8, F32)
    rng = np.random.default_rng(0)
    for scale in (2.0, 4.0):
        edge = F32(scale * a)
        for c, kept in ((edge, True), (np.nextafter(edge, F32(1)), False)):
            v = rng.permutation(base)
            field = np.insert(v, 12, c).reshape(5, 5)   # centre (2, 2) is flat index 12
            keep = _gardner(field * 32, np.zeros((5, 5), F32), 32, iterations=2,
                            mad_scale=scale)
            assert keep[2, 2] == kept, (scale, c)


def test_stage2_runs_iterations_minus_one_times():
    # One grid row, window 5 (1 x 5 neighbourhoods), r = 1, up = 64: floor 1/32. Stage 1 is
    # made a no-op (min_agree_fraction 1/25 -> one agreeing sample, the point itself).
    #   values      0  0  0  0  .25  .5  0
    # Stage-2 pass 1 (every point present):
    #   .25: window [0 0 .25 .5 0] -> median 0, MAD 0 -> tol 1/32 -> rejected
    #   .5 : window [0 .25 .5 0] (clipped) -> median (0+.25)/2 = .125, deviations
    #        [.125 .125 .125 .375] -> MAD .125 -> tol .5 >= .375 -> kept
    #   last 0: window [.25 .5 0] -> median .25, MAD .25 -> kept; the rest: median 0 -> kept
    # Stage-2 pass 2 (.25 now NaN):
    #   .5 : window [0 .5 0] -> median 0, MAD 0 -> tol 1/32 -> rejected
    #   last 0: window [.5 0] -> median .25, MAD .25 -> kept
    # iterations 1 and 2 run one stage-2 pass (max(iter-1, 1)), iterations 3 runs two.
    row = np.array([[0, 0, 0, 0, 0.25, 0.5, 0]], F32)
    zero = np.zeros_like(row)
    one_pass = np.array([[1, 1, 1, 1, 0, 1, 1]], bool)
    two_pass = np.array([[1, 1, 1, 1, 0, 0, 1]], bool)
    for it, exp in ((1, one_pass), (2, one_pass), (3, two_pass)):
        keep = _gardner(row, zero, 1, iterations=it, min_agree_fraction=1 / 25)
        np.testing.assert_array_equal(keep, exp, err_msg=f"iterations={it}")
        keep = _gardner(zero, row, 1, iterations=it, min_agree_fraction=1 / 25)
        np.testing.assert_array_equal(keep, exp, err_msg=f"y axis, iterations={it}")


# --------------------------------------------------------------------------------------------
# Gardner stage 1: |v - centre| < tol counted over the window >= Float32(8/25 * 25) = 8
# (src/outliers.jl:112-128, src/window.jl:594-619)
# --------------------------------------------------------------------------------------------

def _cluster(values):
    """5x5 grid, NaN outside the 3x3 block around the centre; `values` fill the centre first,
    then the block's other cells in column-major order."""
    g = np.full((5, 5), np.nan, F32)
    cells = [(2, 2)] + [(i, j) for j in (1, 2, 3) for i in (1, 2, 3) if (i, j) != (2, 2)]
    for (i, j), v in zip(cells, values):
        g[i, j] = v
    return g


def test_stage1_needs_eight_agreeing_counting_itself():
    # r = 1 (normalised = px), one stage-1 pass and one stage-2 pass.
    kw = dict(iterations=1)
    # 8 zeros: every point counts 8 (itself included) -> kept; stage 2 sees a flat window.
    g = _cluster([0.0] * 8)
    keep = _gardner(g, g, 1, **kw)
    np.testing.assert_array_equal(keep, ~np.isnan(g))
    # 7 zeros: every count is 7 < 8 -> everything rejected.
    g = _cluster([0.0] * 7)
    assert not _gardner(g, g, 1, **kw).any()
    # 7 zeros and one sample at exactly Float32(0.2) from them: the comparison is strict, so
    # no zero counts it -> counts of 7 -> everything rejected.
    t = F32(0.2)
    g = _cluster([0.0] * 7 + [t])
    assert not _gardner(g, np.zeros_like(g), 1, **kw).any()
    # ... one ulp closer and it counts: all 8 pass stage 1. Stage 2 (floor 2/64): the zeros'
    # windows have median 0 and MAD 0, so the zeros stay and the outlying sample goes.
    t2 = np.nextafter(t, F32(0))
    g = _cluster([0.0] * 7 + [t2])
    keep = _gardner(g, np.where(np.isnan(g), np.nan, 0).astype(F32), 1, **kw)
    exp = (g == 0)
    np.testing.assert_array_equal(keep, exp)


# --------------------------------------------------------------------------------------------
# _fill_holes!: >= 2*w^2 / 3 = 6 measured of the 3x3 window, median of each axis with the
# mean of the middle two for an even count, two-phase passes (src/multichip.jl:347-390)
# --------------------------------------------------------------------------------------------

def test_fill_even_count_median_and_six_of_nine():
    nan = np.nan
    # Centre hole of a 3x3 grid, six measured neighbours (two more holes, which have at most
    # five measured neighbours each and so are never filled).
    dx = np.array([[7, 3, nan], [1, nan, 2], [nan, 8, 5]], F32)
    dy = np.array([[-1, 2, nan], [6, nan, -4], [nan, 0, 9]], F32)
    fx, fy, filled = oracle.fill_holes(dx, dy)
    # dx: sorted 1 2 3 5 7 8 -> (3 + 5) / 2 = 4; dy: -4 -1 0 2 6 9 -> (0 + 2) / 2 = 1
    assert fx[1, 1] == 4.0 and fy[1, 1] == 1.0
    assert list(filled) == [1 + 1 * 3]
    assert np.isnan(fx[0, 2]) and np.isnan(fx[2, 0])
    # Eight distinct neighbours: (4 + 5) / 2.
    dx = np.array([[1, 2, 3], [4, nan, 5], [6, 7, 8]], F32)
    fx, fy, _ = oracle.fill_holes(dx, dx * 10)
    assert fx[1, 1] == 4.5 and fy[1, 1] == 45.0
    # Five measured neighbours: not filled.
    dx = np.array([[7, 3, nan], [1, nan, nan], [nan, 8, 5]], F32)
    fx, fy, filled = oracle.fill_holes(dx, dx)
    assert np.isnan(fx[1, 1]) and len(filled) == 0


def test_fill_is_two_phase():
    # 4x4 grid, dx = 10 i + j (0-based), holes at (1,1), (2,1), (1,2); each has exactly six
    # measured neighbours, so all three fill in the first pass from the field as it was:
    #   (1,1): 0 1 2 10 20 22   -> (2 + 10) / 2 = 6
    #   (2,1): 10 20 22 30 31 32 -> (22 + 30) / 2 = 26   (in-place filling would see (1,1) = 6)
    #   (1,2): 1 2 3 13 22 23   -> (3 + 13) / 2 = 8
    i, j = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    dx = (10 * i + j).astype(F32)
    for h in ((1, 1), (2, 1), (1, 2)):
        dx[h] = np.nan
    fx, fy, filled = oracle.fill_holes(dx, -dx)
    assert (fx[1, 1], fx[2, 1], fx[1, 2]) == (6, 26, 8)
    assert (fy[1, 1], fy[2, 1], fy[1, 2]) == (-6, -26, -8)
    assert list(filled) == [5, 6, 9]   # column-major linear indices, in visiting order


# --------------------------------------------------------------------------------------------
# _cell_max_radius!: cells (i-2):(i+1) x (j-2):(j+1) at stride 4 (src/multichip.jl:292-306)
# --------------------------------------------------------------------------------------------

def test_cell_max_radius_margins():
    # 1-based coarse rows/cols 4 and 8 of an 8x8 grid: cells cover rows 2..5 and 6..8.
    r = np.full((8, 8), 6, np.int64)
    r[0, 0] = 99      # row/col 1: in no cell
    r[1, 1] = 50      # (2,2): first row/col of cell (4,4)
    r[4, 6] = 40      # (5,7): last row of cell (4,*), col 7 in cell (*,8)
    r[5, 2] = 30      # (6,3): first row of cell (8,*), col 3 in cell (*,4)
    r[7, 7] = 20      # (8,8): cell (8,8)
    out = oracle.cell_max_radius(r, 4)
    np.testing.assert_array_equal(out, [[50, 40], [30, 20]])
    # stride 3: lo 1, hi 1 -> cells (i-1):(i+1); a 7x7 grid has coarse rows 3 and 6
    r = np.zeros((7, 7), np.int64)
    r[0, 0] = 9       # (1,1): in no cell (cell 3 covers rows and cols 2..4)
    r[1, 1] = 8       # (2,2) -> cell (3,3)
    r[6, 6] = 7       # (7,7) -> cell (6,6) covers 5..7
    np.testing.assert_array_equal(oracle.cell_max_radius(r, 3), [[8, 0], [0, 7]])


# --------------------------------------------------------------------------------------------
# coarse gate: kept / measured-and-searchable >= min fraction (src/multichip.jl:265-279), and
# the all-true fallback for a coarse grid narrower than the filter window (:236-238)
# --------------------------------------------------------------------------------------------

def test_coarse_fraction_gate():
    n = 100
    keep = np.zeros(n, bool)
    keep[0] = True
    meas = np.ones(n, bool)
    rx = np.full(n, 25, np.int64)
    assert oracle.coarse_fraction_ok(keep, meas, rx, 0.01)          # 1/100 is not < 0.01
    assert not oracle.coarse_fraction_ok(keep, meas, rx, 0.0101)
    # a 101st searched-and-measured point tips it below
    k2, m2, r2 = np.append(keep, False), np.append(meas, True), np.append(rx, 25)
    assert not oracle.coarse_fraction_ok(k2, m2, r2, 0.01)
    # ... but not when it was unsearchable (radius 0) or unmeasured: those are not counted
    assert oracle.coarse_fraction_ok(k2, m2, np.append(rx, 0), 0.01)
    assert oracle.coarse_fraction_ok(k2, np.append(meas, False), r2, 0.01)
    # nothing searchable and measured: the level has nothing to go on
    assert not oracle.coarse_fraction_ok(keep, np.zeros(n, bool), rx, 0.0)
    assert not oracle.coarse_fraction_ok(keep, meas, np.zeros(n, np.int64), 0.0)


def test_coarse_fallback_when_grid_narrower_than_window():
    # Flat images: every chip is degenerate, so a coarse pass that runs measures nothing and
    # returns "nothing". At stride 4 a grid of 19 rows has 4 coarse rows < window 5, so the pass
    # does not run and the whole grid is searched; at 20 rows it runs.
    img = np.zeros((200, 200), np.uint8)
    ok = np.ones(img.shape, bool)
    for nr, exp in ((19, "all"), (20, None)):
        xs = 60.0 + 4 * np.arange(20)
        ys = 60.0 + 4 * np.arange(nr)
        X, Y = np.meshgrid(xs, ys)
        m = oracle.coarse_pass(img, img, ok, X, Y, 6, 6, 16, 16, p=oracle.params(chip_size=16))
        if exp == "all":
            assert m is not None and m.shape == X.shape and m.all()
        else:
            assert m is None
    # NoOutlierFilter needs no window: the pass always runs
    m = oracle.coarse_pass(img, img, ok, X[:19], Y[:19], 6, 6, 16, 16,
                           p=oracle.params(chip_size=16, outliers=False))
    assert m is None


# --------------------------------------------------------------------------------------------
# _merge_level!: the first level to measure a point owns it (src/multichip.jl:401-422)
# --------------------------------------------------------------------------------------------

def test_merge_level():
    nan = np.nan
    rdx = np.array([nan, 1, nan, nan, nan], F32)
    rdy = np.array([nan, 2, nan, nan, nan], F32)
    rc = np.array([nan, .5, nan, nan, nan], F32)
    rcs = np.array([0, 32, 0, 0, 0], np.uint16)
    rint = np.array([0, 0, 0, 0, 0], np.uint8)
    level = (np.array([5, 6, nan, 7, 8], F32), np.array([-5, -6, nan, -7, -8], F32),
             np.array([.9, .8, nan, nan, .7], F32))
    oracle.merge_level((rdx, rdy, rc, rcs, rint), level, [3, 1], 64)
    np.testing.assert_array_equal(rdx, [5, 1, nan, 7, 8])
    np.testing.assert_array_equal(rdy, [-5, 2, nan, -7, -8])
    np.testing.assert_array_equal(rc, np.array([.9, .5, nan, nan, .7], F32))
    np.testing.assert_array_equal(rcs, [64, 32, 0, 64, 64])
    # (1) was filled at this level but a finer level owns it: its flag is not touched
    np.testing.assert_array_equal(rint, [0, 0, 0, 1, 0])
