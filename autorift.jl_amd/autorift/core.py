"""Device context: the Python face of one ar_context (one AutoRIFT.jl Cache's device state).

Arrays are numpy arrays in their logical (row, column) shape; they are handed to the library
in column-major order, which is Julia's layout and the ABI's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import AR_F32, AR_NCC, AR_U8, AR_ZNCC, Points, check, lib

_DTYPES = {np.dtype(np.uint8): _lib.AR_U8, np.dtype(np.float32): _lib.AR_F32,
           np.dtype(np.float64): _lib.AR_F64, np.dtype(np.int16): _lib.AR_I16,
           np.dtype(np.complex64): _lib.AR_C64}
_NP = {_lib.AR_U8: np.uint8, _lib.AR_F32: np.float32, _lib.AR_C64: np.complex64}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _fortran(a, dtype=None):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


@dataclass
class DisplacementField:
    """Per-point results of a correlation pass (src/track.jl:46-62)."""
    dx: np.ndarray
    dy: np.ndarray
    correlation: np.ndarray
    searched: np.ndarray

    def nmeasured(self) -> int:
        return int(np.count_nonzero(~np.isnan(self.dx)))


@dataclass
class MultichipResult:
    """Displacement over the full grid, assembled from all chip-size levels
    (src/multichip.jl:74-80)."""
    dx: np.ndarray
    dy: np.ndarray
    correlation: np.ndarray
    chip_size: np.ndarray
    interpolated: np.ndarray

    def nmeasured(self) -> int:
        return int(np.count_nonzero(~np.isnan(self.dx)))


def _points_struct(pts):
    """Build ar_points over the (flattened, column-major) fields of a PointSet."""
    fields = []
    for name, t in (("x", np.float64), ("y", np.float64), ("radius_x", np.int64),
                    ("radius_y", np.int64), ("dx_prior", np.float64), ("dy_prior", np.float64),
                    ("chip_size_x", np.int64), ("chip_size_y", np.int64)):
        v = getattr(pts, name)
        fields.append(np.ascontiguousarray(np.ravel(np.asarray(v, dtype=t), order="F")))
    s = Points(*[_ptr(f) for f in fields])
    return s, fields


class Context:
    """Device state for one image pair: prepared images and masks resident in HBM."""

    def __init__(self, params: _lib.Params | None = None, device: int = 0):
        self._h = C.c_void_p()
        p = params if params is not None else _lib.default_params()
        self.params = p
        check(lib().ar_context_create(C.byref(p), int(device), C.byref(self._h)))
        self.shape = None

    # -- lifecycle -----------------------------------------------------------------------------
    def close(self):
        if self._h:
            lib().ar_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        check(st, self._h)

    # -- images --------------------------------------------------------------------------------
    def set_images(self, reference=None, secondary=None, reference_valid=None,
                   secondary_valid=None):
        """Upload a raw pair and prepare it on the device (init/reinit!, src/api.jl:66-134).
        A None image keeps the slot's current image."""
        ref = None if reference is None else _fortran(reference)
        sec = None if secondary is None else _fortran(secondary)
        some = ref if ref is not None else sec
        if some is None:
            return
        if ref is not None and sec is not None and ref.shape != sec.shape:
            raise _lib.DimensionMismatch(f"reference is {ref.shape} but secondary is {sec.shape}")
        dt = some.dtype
        if dt not in _DTYPES:
            raise _lib.ArgumentError(f"unsupported image element type {dt}")
        if ref is not None and sec is not None and sec.dtype != ref.dtype:
            sec = _fortran(sec, ref.dtype)
        rv = None if reference_valid is None else _fortran(reference_valid, np.uint8)
        sv = None if secondary_valid is None else _fortran(secondary_valid, np.uint8)
        ny, nx = some.shape
        self._check(lib().ar_set_images(self._h, _ptr(ref), _ptr(sec), _ptr(rv), _ptr(sv), ny, nx,
                                        _DTYPES[dt]))
        self.shape = (ny, nx)

    def set_pair(self, reference, secondary, reference_valid=None, secondary_valid=None):
        """Install an already prepared pair, used as is (what track! correlates)."""
        ref = _fortran(reference)
        sec = _fortran(secondary, ref.dtype)
        if ref.shape != sec.shape:
            raise _lib.DimensionMismatch(f"reference is {ref.shape} but secondary is {sec.shape}")
        if ref.dtype not in (np.dtype(np.uint8), np.dtype(np.float32)):
            raise _lib.ArgumentError("a prepared pair is UInt8 or Float32")
        rv = None if reference_valid is None else _fortran(reference_valid, np.uint8)
        sv = None if secondary_valid is None else _fortran(secondary_valid, np.uint8)
        ny, nx = ref.shape
        self._check(lib().ar_set_pair(self._h, _ptr(ref), _ptr(sec), _ptr(rv), _ptr(sv), ny, nx,
                                      _DTYPES[ref.dtype]))
        self.shape = (ny, nx)

    def get_pair(self):
        """(reference, secondary, reference_valid, secondary_valid) as prepared on the device."""
        ny, nx = self.shape
        dt = C.c_int32()
        lib().ar_get_pair(self._h, None, None, None, None, C.byref(dt))
        t = _NP[dt.value]
        r = np.empty((ny, nx), t, order="F")
        s = np.empty((ny, nx), t, order="F")
        rv = np.empty((ny, nx), np.uint8, order="F")
        sv = np.empty((ny, nx), np.uint8, order="F")
        self._check(lib().ar_get_pair(self._h, _ptr(r), _ptr(s), _ptr(rv), _ptr(sv), C.byref(dt)))
        return r, s, rv.astype(bool), sv.astype(bool)

    def quantizer_stats(self, walk=False):
        """The UInt8 quantiser's (n, mean, sd, mode) of the prepared reference and secondary
        (src/preprocess.jl:546-592), as two tuples; with `walk`, each tuple also carries the
        device walk's (runs, elements added one by one) for the sum and for the squared
        deviations."""
        out = (C.c_double * 16)()
        self._check(lib().ar_quantizer_stats(self._h, out))
        v = list(out)
        res = []
        for i in range(2):
            o = v[8 * i:8 * i + 8]
            r = (int(o[0]), o[1], o[2], int(o[3]))
            if walk:
                r += ((int(o[4]), int(o[6])), (int(o[5]), int(o[7])))
            res.append(r)
        return tuple(res)

    def sequential_sum(self, x, method=0):
        """sum(x) added left to right in Float64, computed on the device by the chunk walk
        (method 0: returns (sum, runs walked, elements added one by one)) or by the one-wave
        replay (method 1: non-negative finite terms; returns (sum, 0, 0))."""
        x = np.ascontiguousarray(x, np.float64).ravel()
        out = C.c_double()
        walk = np.zeros(2, np.int64)
        self._check(lib().ar_sequential_sum(self._h, _ptr(x), x.size, method, C.byref(out),
                                            _ptr(walk)))
        return out.value, int(walk[0]), int(walk[1])

    # -- correlation -----------------------------------------------------------------------------
    def track(self, pts, upsampling: int = 1, measure: int | None = None) -> DisplacementField:
        """track (src/track.jl:103-160) over every point of `pts`; `measure` defaults to the
        context's (the first of its similarity measures)."""
        if measure is None:
            measure = self.params.measure
        s, keep = _points_struct(pts)
        shape = np.shape(pts.x)
        n = int(np.prod(shape))
        dx = np.empty(n, np.float32)
        dy = np.empty(n, np.float32)
        cc = np.empty(n, np.float32)
        sr = np.empty(n, np.uint8)
        if n:
            self._check(lib().ar_track(self._h, C.byref(s), n, int(upsampling), int(measure),
                                       _ptr(dx), _ptr(dy), _ptr(cc), _ptr(sr)))
        r = lambda a: a.reshape(shape, order="F")
        return DisplacementField(r(dx), r(dy), r(cc), r(sr).astype(bool))

    def track_device(self, dev_fields, n, upsampling, measure, dev_out):
        """ar_track_device with raw device addresses (ints): dev_fields in PointSet order,
        dev_out = (dx, dy, corr, searched)."""
        s = Points(*[C.c_void_p(int(a)) for a in dev_fields])
        o = [C.c_void_p(int(a)) for a in dev_out]
        self._check(lib().ar_track_device(self._h, C.byref(s), int(n), int(upsampling),
                                          int(measure), *o))

    def run_multichip(self, grid) -> MultichipResult:
        """correlate_multichip (src/multichip.jl:106-137) over a gridded PointSet."""
        s, keep = _points_struct(grid)
        shape = np.shape(grid.x)
        if len(shape) != 2:
            raise _lib.ArgumentError("the multi-chip search needs a gridded (2-D) point set")
        nr, nc = shape
        n = nr * nc
        dx = np.empty(n, np.float32)
        dy = np.empty(n, np.float32)
        cc = np.empty(n, np.float32)
        cs = np.empty(n, np.uint16)
        it = np.empty(n, np.uint8)
        self._check(lib().ar_run_multichip(self._h, C.byref(s), nr, nc, _ptr(dx), _ptr(dy),
                                           _ptr(cc), _ptr(cs), _ptr(it)))
        r = lambda a: a.reshape(shape, order="F")
        return MultichipResult(r(dx), r(dy), r(cc), r(cs), r(it).astype(bool))

    def correlate_point(self, search, chip, rx, ry=None, upsampling=1, measure=AR_ZNCC):
        """correlate! + peak_offset/subpixel_peak for one chip and window: (dx, dy, corr)."""
        ry = rx if ry is None else ry
        chip = _fortran(chip)
        search = _fortran(search, chip.dtype)
        ch, cw = chip.shape
        if search.shape != (ch + 2 * ry - 1, cw + 2 * rx - 1):
            raise _lib.DimensionMismatch(
                f"search window is {search.shape} but a {ch}x{cw} chip with radius ({rx}, {ry}) "
                f"requires {(ch + 2 * ry - 1, cw + 2 * rx - 1)}")
        dx, dy, v = C.c_double(), C.c_double(), C.c_float()
        self._check(lib().ar_correlate_point(self._h, _ptr(search), _ptr(chip),
                                             _DTYPES[chip.dtype], ch, cw, rx, ry,
                                             int(upsampling), int(measure), C.byref(dx),
                                             C.byref(dy), C.byref(v)))
        return dx.value, dy.value, v.value

    def last_kernel(self):
        ms, path = C.c_double(), C.c_int32()
        lib().ar_last_kernel(self._h, C.byref(ms), C.byref(path))
        return ms.value, path.value

    def last_timing(self):
        """(box images ms, correlation kernel ms, refinement kernel ms) of the last pass."""
        b, t, r = C.c_double(), C.c_double(), C.c_double()
        lib().ar_last_timing(self._h, C.byref(b), C.byref(t), C.byref(r))
        return b.value, t.value, r.value
