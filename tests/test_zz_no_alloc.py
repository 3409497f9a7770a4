"""GPU test of ar_correlate_point_noalloc: ar_correlate_point with no device allocation (a
library-owned device array instead), called through the C ABI. Each result must equal
ar_correlate_point's bit for bit; two calls in a row, the second on a smaller input; sizes past the
array are refused.
"""
import ctypes as C

import numpy as np
import pytest

from _util import chip_and_window, synthetic_texture

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")
L = autorift._lib


def _call(fn, ctx, search, chip, r, up, measure):
    search = np.asfortranarray(search)
    chip = np.asfortranarray(chip)
    dt = L.AR_U8 if chip.dtype == np.uint8 else L.AR_F32
    dx, dy, corr = C.c_double(), C.c_double(), C.c_float()
    st = fn(ctx._h, search.ctypes.data_as(C.c_void_p), chip.ctypes.data_as(C.c_void_p), dt,
            chip.shape[0], chip.shape[1], r, r, up, measure, C.byref(dx), C.byref(dy),
            C.byref(corr))
    return st, (dx.value, dy.value, corr.value)


def _case(cs, r, dtype, seed):
    img = synthetic_texture((cs + 2 * r + 40, cs + 2 * r + 40), seed=seed)
    if dtype == np.uint8:
        img = np.rint(img * 255).astype(np.uint8)
    else:
        img = (img * 255).astype(np.float32)
    c = (img.shape[0] // 2, img.shape[1] // 2)
    chip, win = chip_and_window(img, c, cs, r, shift=(2, -3))
    return win, chip


def _bits(v):
    return np.array(v[:2], np.float64).view(np.uint64).tolist() + \
        [np.float32(v[2]).view(np.uint32).item()]


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_noalloc_equals_correlate_point(dtype):
    lib = L.lib()
    with autorift.Context() as ctx:
        # two calls in a row, the second smaller; then the large one again
        for k, (cs, r) in enumerate(((48, 25), (16, 6), (32, 25), (64, 20))):
            win, chip = _case(cs, r, dtype, seed=k)
            for up in (1, 64):
                for measure in (L.AR_ZNCC, L.AR_NCC):
                    print(f"noalloc {np.dtype(dtype)} chip {cs} r {r} up {up} m {measure}",
                          flush=True)
                    st0, exp = _call(lib.ar_correlate_point, ctx, win, chip, r, up, measure)
                    st1, got = _call(lib.ar_correlate_point_noalloc, ctx, win, chip, r, up,
                                     measure)
                    assert st0 == 0 and st1 == 0, lib.ar_last_error(ctx._h)
                    assert _bits(got) == _bits(exp), (got, exp)
        # a flat chip is degenerate: NaN, as ar_correlate_point reports
        win, chip = _case(16, 6, dtype, seed=9)
        chip = np.full_like(chip, 7)
        st, got = _call(lib.ar_correlate_point_noalloc, ctx, win, chip, 6, 64, L.AR_ZNCC)
        assert st == 0 and np.isnan(got[0]) and np.isnan(got[1])


def test_noalloc_refuses_oversized_and_keeps_the_pair():
    lib = L.lib()
    with autorift.Context() as ctx:
        ref = (synthetic_texture((200, 200), seed=3) * 255).astype(np.float32)
        ctx.set_pair(ref, ref)
        before = ctx.get_pair()[0].copy()
        win, chip = _case(256, 10, np.float32, seed=4)
        st, _ = _call(lib.ar_correlate_point_noalloc, ctx, win, chip, 10, 1, L.AR_ZNCC)
        assert st == L.AR_ERR_UNSUPPORTED
        assert b"too large" in lib.ar_last_error(ctx._h)
        win, chip = _case(32, 12, np.float32, seed=5)
        st, _ = _call(lib.ar_correlate_point_noalloc, ctx, win, chip, 12, 16, L.AR_ZNCC)
        assert st == 0
        np.testing.assert_array_equal(ctx.get_pair()[0], before)
