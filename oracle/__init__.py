"""ctypes front end of the CPU oracle (test infrastructure only).

Only the test suite, the smoke check and the CPU-baseline leg of the benchmark import this
package, and only as the checker. Arrays are numpy arrays in their logical (row, col) shape;
they are handed to C in column-major (Fortran) order, which is Julia's layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

U8, F32, F64, I16, C64 = 0, 1, 2, 3, 4
ZNCC, NCC, COHERENCE = 0, 1, 2
NUM_DIRECT, NUM_FFT = 0, 1
ERR_MEASURE, ERR_ROTATION = -2, -3

_DT = {np.dtype(np.uint8): U8, np.dtype(np.float32): F32, np.dtype(np.float64): F64,
       np.dtype(np.int16): I16, np.dtype(np.complex64): C64}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_next_fft_size.restype = C.c_int64
        _lib.orc_next_fft_size.argtypes = [C.c_int64]
        _lib.orc_fill_holes.restype = C.c_int64
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f(a, dtype=None):
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a)


class _Points(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("x", "y", "radius_x", "radius_y", "dx_prior", "dy_prior", "chip_size_x",
                 "chip_size_y")]


def _points(x, y, rx, ry, pdx, pdy, csx, csy):
    keep = [np.ascontiguousarray(np.ravel(np.asarray(v, dtype=t), order="F"))
            for v, t in ((x, np.float64), (y, np.float64), (rx, np.int64), (ry, np.int64),
                         (pdx, np.float64), (pdy, np.float64), (csx, np.int64),
                         (csy, np.int64))]
    s = _Points(*[_p(k) for k in keep])
    return s, keep


class _Gardner(C.Structure):
    _fields_ = [("window", C.c_int64), ("iterations", C.c_int64),
                ("min_agree_fraction", C.c_double), ("agree_tolerance", C.c_double),
                ("mad_scale", C.c_double)]


class _Params(C.Structure):
    _fields_ = [("chip_size_base", C.c_int64), ("chip_size_min", C.c_int64),
                ("chip_size_max", C.c_int64), ("chip_aspect", C.c_double),
                ("min_search_radius", C.c_int64), ("coarse_stride", C.c_int64),
                ("coarse_buffer", C.c_int64), ("min_coarse_valid_fraction", C.c_double),
                ("fill_window", C.c_int64), ("upsampling", C.c_int64), ("measure", C.c_int),
                ("outliers_on", C.c_int), ("gardner", _Gardner), ("nrot", C.c_int),
                ("angles", C.c_double * 8), ("nmeas", C.c_int), ("measures", C.c_int * 8)]


def next_fft_size(n: int) -> int:
    return int(lib().orc_next_fft_size(int(n)))


def correlate(search, chip, rx, ry=None, measure=ZNCC, num_mode=NUM_DIRECT):
    """correlate! (src/correlate.jl:528-557): returns (surface, ok)."""
    ry = rx if ry is None else ry
    chip = _f(chip)
    search = _f(search, chip.dtype)
    ch, cw = chip.shape
    assert search.shape == (ch + 2 * ry - 1, cw + 2 * rx - 1)
    out = np.zeros((2 * ry, 2 * rx), dtype=np.float32, order="F")
    ok = lib().orc_correlate(_p(search), C.c_int64(search.shape[0]), _p(chip),
                             C.c_int64(ch), _DT[chip.dtype], C.c_int64(ch), C.c_int64(cw),
                             C.c_int64(rx), C.c_int64(ry), measure, num_mode, _p(out))
    return out, bool(ok)


def integral_both(a):
    a = _f(a)
    m, n = a.shape
    S = np.zeros((m + 1, n + 1), order="F")
    S2 = np.zeros((m + 1, n + 1), order="F")
    lib().orc_integral_both(_p(a), C.c_int64(m), _DT[a.dtype], C.c_int64(m), C.c_int64(n),
                            _p(S), _p(S2))
    return S, S2


def peak_index(surface):
    """1-based (row, col) -- src/peak.jl:32-61."""
    s = _f(surface, np.float32)
    r, c = C.c_int64(), C.c_int64()
    lib().orc_peak_index(_p(s), C.c_int64(s.shape[0]), C.c_int64(s.shape[1]), C.byref(r),
                         C.byref(c))
    return int(r.value), int(c.value)


def pyrup(src):
    s = _f(src, np.float32)
    out = np.zeros((2 * s.shape[0], 2 * s.shape[1]), dtype=np.float32, order="F")
    lib().orc_pyrup(_p(out), _p(s), C.c_int64(s.shape[0]), C.c_int64(s.shape[1]))
    return out


def subpixel_peak(surface, rx, ry, up):
    s = _f(surface, np.float32)
    dx, dy, v = C.c_double(), C.c_double(), C.c_float()
    lib().orc_subpixel_peak(_p(s), C.c_int64(s.shape[0]), C.c_int64(s.shape[1]),
                            C.c_int64(rx), C.c_int64(ry), C.c_int64(up), C.byref(dx),
                            C.byref(dy), C.byref(v))
    return dx.value, dy.value, v.value


def _angles(rotation):
    a = [] if rotation is None else [float(v) for v in rotation]
    assert len(a) <= 8
    return len(a), (C.c_double * 8)(*(a + [0.0] * (8 - len(a))))


def track(ref, sec, ok_mask, x, y, rx, ry, csx, csy, pdx=0.0, pdy=0.0, up=1, measure=ZNCC,
          num_mode=NUM_DIRECT, nthreads=1, rotation=None):
    """track (src/track.jl:103-160) over a flat or gridded point set.

    Point fields broadcast to the shape of ``x``. Returns (dx, dy, corr, searched) shaped
    like ``x``. ``rotation``: the chip rotations of a RotationSearch (degrees, ``about``
    already subtracted), or None.
    """
    ref = _f(ref)
    sec = _f(sec, ref.dtype)
    ok = _f(ok_mask, np.uint8)
    shape = np.shape(x)
    b = lambda v, t: np.broadcast_to(np.asarray(v, dtype=t), shape)
    pts, keep = _points(b(x, np.float64), b(y, np.float64), b(rx, np.int64), b(ry, np.int64),
                        b(pdx, np.float64), b(pdy, np.float64), b(csx, np.int64),
                        b(csy, np.int64))
    n = int(np.prod(shape))
    dx = np.empty(n, np.float32)
    dy = np.empty(n, np.float32)
    corr = np.empty(n, np.float32)
    srch = np.empty(n, np.uint8)
    nrot, ang = _angles(rotation)
    rc = lib().orc_track(_p(ref), _p(sec), _p(ok), _DT[ref.dtype], C.c_int64(ref.shape[0]),
                         C.c_int64(ref.shape[1]), C.byref(pts), C.c_int64(n), C.c_int64(up),
                         measure, num_mode, nthreads, nrot, ang, _p(dx), _p(dy), _p(corr),
                         _p(srch))
    if rc == ERR_MEASURE:
        raise ValueError("Coherence needs complex images and ZNCC / NCC real ones")
    if rc == ERR_ROTATION:
        raise ValueError("rotation search needs real images")
    r = lambda a: a.reshape(shape, order="F")
    return r(dx), r(dy), r(corr), r(srch).astype(bool)


def windowmean(a, w, hasnan=-1):
    a = _f(a, np.float32)
    wx, wy = (w, w) if np.isscalar(w) else w
    out = np.empty(a.shape, np.float32, order="F")
    lib().orc_windowmean(_p(a), C.c_int64(a.shape[0]), C.c_int64(a.shape[1]), C.c_int64(wx),
                         C.c_int64(wy), hasnan, _p(out))
    return out


def window_extremum(a, w, is_max):
    a = _f(a, np.float32)
    wx, wy = (w, w) if np.isscalar(w) else w
    out = np.empty(a.shape, np.float32, order="F")
    lib().orc_window_extremum(_p(a), C.c_int64(a.shape[0]), C.c_int64(a.shape[1]),
                              C.c_int64(wx), C.c_int64(wy), int(bool(is_max)), _p(out))
    return out


def highpass_filtered(img, mask, w=5):
    img = _f(img)
    mask = _f(mask, np.uint8)
    out = np.empty(img.shape, np.float32, order="F")
    vm = np.empty(img.shape, np.uint8, order="F")
    lib().orc_highpass_filtered(_p(img), _p(mask), _DT[img.dtype], C.c_int64(img.shape[0]),
                                C.c_int64(img.shape[1]), C.c_int64(w), _p(out), _p(vm))
    return out, vm.astype(bool)


def wallis_filtered(img, mask, w=5, min_std=0.0):
    """preprocess(img, mask, Wallis(w, min_std)) (src/preprocess.jl:158-159, 333-351, 416-438):
    (filtered Float32 image, eroded validity)."""
    img = _f(img)
    mask = _f(mask, np.uint8)
    out = np.empty(img.shape, np.float32, order="F")
    vm = np.empty(img.shape, np.uint8, order="F")
    lib().orc_wallis_filtered(_p(img), _p(mask), _DT[img.dtype], C.c_int64(img.shape[0]),
                              C.c_int64(img.shape[1]), C.c_int64(w), C.c_double(min_std),
                              _p(out), _p(vm))
    return out, vm.astype(bool)


def to_uint8(img, mask):
    img = _f(img, np.float32)
    mask = _f(mask, np.uint8)
    out = np.empty(img.shape, np.uint8, order="F")
    lib().orc_to_uint8(_p(img), _p(mask), C.c_int64(img.shape[0]), C.c_int64(img.shape[1]),
                       _p(out))
    return out


def quant_stats(img, mask):
    """(n, mean, sd) of the UInt8 quantiser (src/preprocess.jl:547-570): left-to-right Float64
    sums over the valid finite values, in column-major order."""
    v = _f(img, np.float64)
    mask = _f(mask, np.uint8)
    out = np.empty(3, np.float64)
    lib().orc_quant_stats(_p(v), _p(mask), C.c_int64(v.size), _p(out))
    return int(out[0]), float(out[1]), float(out[2])


def count_agreeing(a, w, tol):
    a = _f(a, np.float32)
    out = np.empty(a.shape, np.float32, order="F")
    lib().orc_count_agreeing(_p(a), C.c_int64(a.shape[0]), C.c_int64(a.shape[1]), C.c_int64(w),
                             C.c_float(tol), _p(out))
    return out


def windowmedmad(a, w):
    a = _f(a, np.float32)
    med = np.empty(a.shape, np.float32, order="F")
    mad = np.empty(a.shape, np.float32, order="F")
    lib().orc_windowmedmad(_p(a), C.c_int64(a.shape[0]), C.c_int64(a.shape[1]), C.c_int64(w),
                           _p(med), _p(mad))
    return med, mad


def gardner(window=5, iterations=3, min_agree_fraction=8 / 25, agree_tolerance=0.2,
            mad_scale=4.0):
    return _Gardner(window, iterations, min_agree_fraction, agree_tolerance, mad_scale)


def reject_gardner(dx, dy, rx, ry, valid, up, g=None):
    g = g or gardner()
    dx = _f(dx, np.float32)
    dy = _f(dy, np.float32)
    rx = _f(rx, np.int64)
    ry = _f(ry, np.int64)
    valid = _f(valid, np.uint8)
    keep = np.empty(dx.shape, np.uint8, order="F")
    lib().orc_reject_gardner(_p(dx), _p(dy), _p(rx), _p(ry), _p(valid), C.c_int64(dx.shape[0]),
                             C.c_int64(dx.shape[1]), C.c_int64(up), C.byref(g), _p(keep))
    return keep.astype(bool)


def sqdist_transform(mask):
    m = _f(mask, np.uint8)
    d2 = np.empty(m.shape, np.float64, order="F")
    lib().orc_sqdist_transform(_p(m), C.c_int64(m.shape[0]), C.c_int64(m.shape[1]), _p(d2))
    return d2


def dilate_within(mask, radius):
    m = _f(mask, np.uint8)
    out = np.empty(m.shape, np.uint8, order="F")
    lib().orc_dilate_within(_p(m), C.c_int64(m.shape[0]), C.c_int64(m.shape[1]),
                            C.c_double(radius), _p(out))
    return out.astype(bool)


def resample_nearest(a, shape):
    a = _f(a, np.float32)
    out = np.empty(shape, np.float32, order="F")
    lib().orc_resample_nearest(_p(a), C.c_int64(a.shape[0]), C.c_int64(a.shape[1]),
                               C.c_int64(shape[0]), C.c_int64(shape[1]), _p(out))
    return out


def fill_holes(dx, dy, w=3):
    dx = np.array(_f(dx, np.float32), order="F")
    dy = np.array(_f(dy, np.float32), order="F")
    filled = np.empty(dx.size, np.int64)
    n = lib().orc_fill_holes(_p(dx), _p(dy), C.c_int64(dx.shape[0]), C.c_int64(dx.shape[1]),
                             C.c_int64(w), _p(filled))
    return dx, dy, filled[:n].copy()


def params(chip_size=32, chip_size_min=None, chip_size_max=None, chip_aspect=1.0,
           min_search_radius=6, coarse_stride=4, coarse_buffer=8,
           min_coarse_valid_fraction=0.01, fill_window=3, upsampling=64, measure=ZNCC,
           outliers=True, g=None, rotation=None, measures=None):
    """``measures``: one measure per chip-size level, the last repeating (a similarity
    tuple), or None for ``measure`` everywhere."""
    cmin = chip_size if chip_size_min is None else chip_size_min
    cmax = 4 * chip_size if chip_size_max is None else chip_size_max
    nrot, ang = _angles(rotation)
    ms = [] if measures is None else [int(m) for m in measures]
    assert len(ms) <= 8
    return _Params(chip_size, cmin, cmax, chip_aspect, min_search_radius, coarse_stride,
                   coarse_buffer, min_coarse_valid_fraction, fill_window, upsampling, measure,
                   int(outliers), g or gardner(), nrot, ang, len(ms),
                   (C.c_int * 8)(*(ms + [0] * (8 - len(ms)))))


def _c64(z):
    return np.asfortranarray(np.asarray(z, dtype=np.complex64))


def ramp_phase(z, mask=None):
    """ramp_phase (src/preprocess.jl:233-245): (phase_x, phase_y) in radians per pixel."""
    z = _c64(z)
    m = _f(np.ones(z.shape, bool) if mask is None else mask, np.uint8)
    px, py = C.c_double(), C.c_double()
    lib().orc_ramp_phase(_p(z), _p(m), C.c_int64(z.shape[0]), C.c_int64(z.shape[1]),
                         C.byref(px), C.byref(py))
    return px.value, py.value


def deramp(z, mask=None, axis="both"):
    """deramp (src/preprocess.jl:184-231): ComplexF32 image with its linear phase ramp
    removed."""
    z = _c64(z)
    m = _f(np.ones(z.shape, bool) if mask is None else mask, np.uint8)
    out = np.empty(z.shape, np.complex64, order="F")
    lib().orc_deramp(_p(z), _p(m), C.c_int64(z.shape[0]), C.c_int64(z.shape[1]),
                     {"both": 0, "x": 1, "y": 2}[axis], _p(out))
    return out


def rotate_chip(chip, deg):
    """_rotate_chip (src/track.jl:382-409): Float32 chip rotated by ``deg`` degrees."""
    chip = _f(chip)
    out = np.empty(chip.shape, np.float32, order="F")
    lib().orc_rotate_chip(_p(chip), C.c_int64(chip.shape[0]), _DT[chip.dtype],
                          C.c_int64(chip.shape[0]), C.c_int64(chip.shape[1]), C.c_double(deg),
                          _p(out))
    return out


def multichip(ref, sec, ok_mask, x, y, rx, ry, pdx=0.0, pdy=0.0, p=None, num_mode=NUM_DIRECT,
              nthreads=1):
    """correlate_multichip (src/multichip.jl:106-137) over a gridded point set."""
    p = p or params()
    ref = _f(ref)
    sec = _f(sec, ref.dtype)
    ok = _f(ok_mask, np.uint8)
    shape = np.shape(x)
    assert len(shape) == 2
    b = lambda v, t: np.broadcast_to(np.asarray(v, dtype=t), shape)
    pts, keep = _points(b(x, np.float64), b(y, np.float64), b(rx, np.int64), b(ry, np.int64),
                        b(pdx, np.float64), b(pdy, np.float64), b(0, np.int64), b(0, np.int64))
    n = int(np.prod(shape))
    dx = np.empty(n, np.float32)
    dy = np.empty(n, np.float32)
    corr = np.empty(n, np.float32)
    cs = np.empty(n, np.uint16)
    interp = np.empty(n, np.uint8)
    lib().orc_multichip(_p(ref), _p(sec), _p(ok), _DT[ref.dtype], C.c_int64(ref.shape[0]),
                        C.c_int64(ref.shape[1]), C.byref(pts), C.c_int64(shape[0]),
                        C.c_int64(shape[1]), C.byref(p), num_mode, nthreads, _p(dx), _p(dy),
                        _p(corr), _p(cs), _p(interp))
    r = lambda a: a.reshape(shape, order="F")
    return r(dx), r(dy), r(corr), r(cs), r(interp).astype(bool)


def cell_max_radius(radius, stride):
    """_cell_max_radius! (src/multichip.jl:292-306) over an nr x nc radius grid."""
    r = _f(radius, np.int64)
    nr, nc = r.shape
    out = np.zeros((nr // stride, nc // stride), np.int64, order="F")
    lib().orc_cell_max_radius(_p(r), C.c_int64(nr), C.c_int64(nc), C.c_int64(stride), _p(out))
    return out


def coarse_fraction_ok(keep, measured, radius_x, min_fraction=0.01):
    """The coarse pass's coherence gate (src/multichip.jl:265-279)."""
    k = np.ascontiguousarray(np.ravel(keep, order="F"), np.uint8)
    m = np.ascontiguousarray(np.ravel(measured, order="F"), np.uint8)
    r = np.ascontiguousarray(np.ravel(radius_x, order="F"), np.int64)
    lib().orc_coarse_fraction_ok.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                             C.c_double]
    return bool(lib().orc_coarse_fraction_ok(_p(k), _p(m), _p(r), k.size, float(min_fraction)))


def coarse_pass(ref, sec, ok_mask, x, y, rx, ry, csx, csy, pdx=0.0, pdy=0.0, p=None,
                num_mode=NUM_DIRECT):
    """_coarse_pass (src/multichip.jl:224-288) over a gridded level point set: the fine-pass
    mask, or None when the level has nothing coherent."""
    p = p or params()
    ref = _f(ref)
    sec = _f(sec, ref.dtype)
    ok = _f(ok_mask, np.uint8)
    shape = np.shape(x)
    b = lambda v, t: np.broadcast_to(np.asarray(v, dtype=t), shape)
    pts, keep = _points(b(x, np.float64), b(y, np.float64), b(rx, np.int64), b(ry, np.int64),
                        b(pdx, np.float64), b(pdy, np.float64), b(csx, np.int64),
                        b(csy, np.int64))
    mask = np.zeros(shape, np.uint8, order="F")
    r = lib().orc_coarse_pass(_p(ref), _p(sec), _p(ok), _DT[ref.dtype], C.c_int64(ref.shape[0]),
                              C.c_int64(ref.shape[1]), C.byref(pts), C.c_int64(shape[0]),
                              C.c_int64(shape[1]), C.c_int64(csx), C.c_int64(csy), C.byref(p),
                              num_mode, 1, _p(mask))
    return mask.astype(bool) if r else None


def merge_level(result, level, filled, chip_size):
    """_merge_level! (src/multichip.jl:401-422). ``result`` = (dx, dy, corr, chip_size,
    interpolated) arrays, updated in place (Fortran order); ``level`` = (dx, dy, corr)."""
    rdx, rdy, rc, rcs, rint = result
    fdx, fdy, fc = (np.asfortranarray(a, np.float32) for a in level)
    f = np.ascontiguousarray(filled, np.int64)
    scratch = np.zeros(rdx.size, np.uint8)
    lib().orc_merge_level(_p(rdx), _p(rdy), _p(rc), _p(rcs), _p(rint), _p(fdx), _p(fdy), _p(fc),
                          _p(f), C.c_int64(f.size), C.c_int64(rdx.size), C.c_int64(chip_size),
                          _p(scratch))
