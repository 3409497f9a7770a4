"""Shared test helpers: golden-fixture loading and seeded synthetic scenes."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    """Arrays of a golden case in their logical (row, col) shape, plus its params.

    Files are C-order, as written by numpy; AutoRIFT.jl's loader permutes them into
    column-major (test/utils.jl:23-63), which is the same logical array.
    """
    d = os.path.join(GOLDEN, name)
    meta = json.load(open(os.path.join(d, "meta.json")))
    arrays = {}
    for key, spec in meta["arrays"].items():
        a = np.fromfile(os.path.join(d, key + ".bin"), dtype=np.dtype(spec["dtype"]))
        arrays[key] = a.reshape(spec["shape"])
    return arrays, meta["params"]


def box3(a):
    """3x3 mean over in-bounds neighbours (test/utils.jl:212-230)."""
    nr, nc = a.shape
    p = np.pad(a, 1)
    ones = np.pad(np.ones_like(a), 1)
    s = sum(p[1 + di:1 + di + nr, 1 + dj:1 + dj + nc] for di in (-1, 0, 1) for dj in (-1, 0, 1))
    n = sum(ones[1 + di:1 + di + nr, 1 + dj:1 + dj + nc] for di in (-1, 0, 1) for dj in (-1, 0, 1))
    return s / n


def synthetic_texture(shape, seed=0):
    """Blurred white noise scaled to [0, 1] (the shape of test/utils.jl:174-186, numpy RNG)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    for _ in range(3):
        a = box3(a)
    a -= a.min()
    a /= a.max()
    return a


def shifted_pair(n, shift, seed=0, dtype=np.float32):
    """(reference, secondary) with secondary = reference translated by shift=(dx, dy).

    Integer shifts are exact; fractional ones use bilinear sampling with clamped edges
    (test/utils.jl:254-264, 302-321).
    """
    dx, dy = shift
    pad = int(np.ceil(max(abs(dx), abs(dy)))) + 2
    big = synthetic_texture((n + 2 * pad, n + 2 * pad), seed)
    nr, nc = big.shape
    I, J = np.meshgrid(np.arange(1, nr + 1), np.arange(1, nc + 1), indexing="ij")
    y, x = I - dy, J - dx
    i0, j0 = np.floor(y).astype(int), np.floor(x).astype(int)
    fy, fx = y - i0, x - j0
    out = np.zeros_like(big)
    for di, wi in ((0, 1 - fy), (1, fy)):
        for dj, wj in ((0, 1 - fx), (1, fx)):
            ii = np.clip(i0 + di, 1, nr) - 1
            jj = np.clip(j0 + dj, 1, nc) - 1
            out += wi * wj * big[ii, jj]
    ref = big[pad:pad + n, pad:pad + n]
    sec = out[pad:pad + n, pad:pad + n]
    if dtype == np.uint8:
        conv = lambda a: np.rint(np.clip(a, 0, 1) * 255).astype(np.uint8)
    else:
        conv = lambda a: a.astype(dtype)
    return conv(ref), conv(sec)


def chip_and_window(img, centre, chip_size, radius, shift=(0, 0), window_from=None):
    """1-based centre (row, col); test/utils.jl:288-299."""
    window_from = img if window_from is None else window_from
    cy, cx = centre
    sy, sx = shift
    h = chip_size // 2
    chip = img[cy - h + sy - 1:cy + h - 1 + sy, cx - h + sx - 1:cx + h - 1 + sx]
    window = window_from[cy - h - radius - 1:cy + h + radius - 2, cx - h - radius - 1:cx + h + radius - 2]
    return chip, window
