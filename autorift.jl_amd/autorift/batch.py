"""Batched passes: many prepared pairs resident on the device and one tracking pass over points
of all of them (the library's ar_batch). Each point names its pair; chip sizes and search radii
may differ from point to point. There is no counterpart in the reference, which tracks one pair
per call (src/track.jl:103-160); per point the work is that of `track`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params, lib
from .api import params as _params
from .core import _DTYPES, DisplacementField, _fortran, _points_struct, _ptr


class Batch:
    """Resident UInt8-prepared pairs on `device` (the parameters must quantise to UInt8).

    Chips whose area is a power of two give results bit-equal to a pass over their pair alone.
    Other chip areas take the integer matrix-core tracker with the unrounded mean-removed chip
    (integer peaks agree with the exact pass to the batched form's tolerance), or, with
    `exact=True`, the direct Float64 tracker, bit-equal as well."""

    def __init__(self, p: Params | None = None, device: int = 0, exact: bool = False):
        self.params = p if p is not None else _params()
        self._h = C.c_void_p()
        st = lib().ar_batch_create(C.byref(self.params), int(device), C.byref(self._h))
        if st != _lib.AR_OK:
            _lib.check(st, None)
        self.shapes = []
        self.set_exact(exact)

    def set_exact(self, exact: bool):
        """Chips whose area is not a power of two: the direct Float64 tracker (True) or the
        integer tracker with the unrounded mean-removed chip (False)."""
        self.exact = bool(exact)
        self._check(lib().ar_batch_set_exact(self._h, int(self.exact)))

    def close(self):
        if self._h:
            lib().ar_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.shapes)

    def _check(self, st):
        if st == _lib.AR_OK:
            return
        msg = lib().ar_batch_last_error(self._h).decode(errors="replace")
        if st == _lib.AR_ERR_INVALID_ARGUMENT:
            raise _lib.ArgumentError(msg)
        if st == _lib.AR_ERR_DIMENSION_MISMATCH:
            raise _lib.DimensionMismatch(msg)
        raise _lib.AutoRIFTError(st, msg)

    def add_pair(self, reference, secondary, reference_valid=None, secondary_valid=None) -> int:
        """Prepare a raw pair (as init does) and keep it; returns its pair index."""
        ref = _fortran(reference)
        sec = _fortran(secondary, ref.dtype)
        if ref.shape != sec.shape:
            raise _lib.DimensionMismatch(f"reference is {ref.shape} but secondary is {sec.shape}")
        if ref.dtype not in _DTYPES:
            raise _lib.ArgumentError(f"unsupported image element type {ref.dtype}")
        rv = None if reference_valid is None else _fortran(reference_valid, np.uint8)
        sv = None if secondary_valid is None else _fortran(secondary_valid, np.uint8)
        k = C.c_int32()
        ny, nx = ref.shape
        self._check(lib().ar_batch_add_pair(self._h, _ptr(ref), _ptr(sec), _ptr(rv), _ptr(sv),
                                            ny, nx, _DTYPES[ref.dtype], C.byref(k)))
        self.shapes.append((ny, nx))
        return k.value

    def add_pair_device(self, ref_ptr: int, sec_ptr: int, ny: int, nx: int,
                        dtype=_lib.AR_F32) -> int:
        """As add_pair with the raw images already in device memory (column-major)."""
        k = C.c_int32()
        self._check(lib().ar_batch_add_pair(self._h, C.c_void_p(ref_ptr), C.c_void_p(sec_ptr),
                                            None, None, ny, nx, dtype, C.byref(k)))
        self.shapes.append((ny, nx))
        return k.value

    def track(self, pts, pair, upsampling: int | None = None, measure: int | None = None):
        """One pass over every point of `pts` (any shape); pair[i] is point i's pair index."""
        up = self.params.upsampling if upsampling is None else int(upsampling)
        m = self.params.measure if measure is None else int(measure)
        s, keep = _points_struct(pts)
        shape = np.shape(pts.x)
        n = int(np.prod(shape))
        pr = np.ascontiguousarray(np.ravel(np.asarray(pair), order="F"), dtype=np.int32)
        if pr.size != n:
            raise _lib.DimensionMismatch(f"{pr.size} pair indices for {n} points")
        dx = np.empty(n, np.float32)
        dy = np.empty(n, np.float32)
        cc = np.empty(n, np.float32)
        sr = np.empty(n, np.uint8)
        self._check(lib().ar_batch_track(self._h, C.byref(s), _ptr(pr), n, up, m, _ptr(dx),
                                         _ptr(dy), _ptr(cc), _ptr(sr)))
        r = lambda a: a.reshape(shape, order="F")
        return DisplacementField(r(dx), r(dy), r(cc), r(sr).astype(bool))

    def track_device(self, fields, pair_ptr: int, n: int, upsampling: int, measure: int, out,
                     graph: bool = False, split: int | None = None):
        """Device pointers: fields = (x, y, radius_x, radius_y, dx_prior, dy_prior, chip_size_x,
        chip_size_y), out = (dx, dy, correlation, searched). graph: capture the pass and replay
        it while the buffers (and the points' radii) stay the same. split: radius classes R1 or
        (R1, R2) instead of the tuner's choice."""
        s = _lib.Points(*[C.c_void_p(f) for f in fields])
        flags = (_lib.AR_BATCH_GRAPH if graph else 0)
        if split is not None:     # R1, or (R1, R2): radius classes, one launch each
            r1, r2 = (split, 0) if np.isscalar(split) else split
            flags |= _lib.AR_BATCH_SPLIT_SET | (int(r1) << 8) | (int(r2) << 16)
        self._check(lib().ar_batch_track_device(self._h, C.byref(s), C.c_void_p(pair_ptr), n,
                                                upsampling, measure,
                                                *[C.c_void_p(o) for o in out], flags))

    def info(self) -> dict:
        ms = C.c_double()
        v = (C.c_int32 * 4)()
        lib().ar_batch_info(self._h, C.byref(ms), v)
        return {"ms": ms.value, "pairs": v[0], "split": (v[1] & 0xff, v[1] >> 8),
                "launches": v[2], "replayed": bool(v[3])}


def tune_cache_path(device: str) -> str | None:
    """The tuning cache file for a device name such as "gfx950-256cu" (None when caching is off
    or no cache directory is usable)."""
    buf = C.create_string_buffer(4096)
    n = lib().ar_tune_cache_path(device.encode(), buf, len(buf))
    return buf.value.decode() if n > 0 else None


def tune_cache_lookup(device: str, key: str) -> int | None:
    v = C.c_int32()
    return v.value if lib().ar_tune_cache_lookup(device.encode(), key.encode(), C.byref(v)) else None


def tune_cache_record(device: str, key: str, value: int) -> bool:
    """Records key = value; False when the write was dropped (it never raises)."""
    return bool(lib().ar_tune_cache_record(device.encode(), key.encode(), int(value)))
