"""Serving: many pairs of different sizes in flight on one GPU (the library's ar_server).

Each pair runs as `autorift(reference, secondary)` does -- preparation, then the multi-chip
search over the grid built from the parameters (src/api.jl:210-274, 339-347) -- on one of the
server's slots. Slots are native worker threads with their own device buffers and HIP stream;
the library admits queued pairs to idle slots by device-memory footprint (FIFO) and the pairs on
different slots overlap on the device. There is no counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params, lib
from .api import _build_grid, params as _params
from .core import _DTYPES, MultichipResult, _fortran, _points_struct, _ptr


class ServerError(_lib.AutoRIFTError):
    pass


class Server:
    """`slots` concurrent pairs on `device`; `hbm_budget` bytes of device memory (None: 90 % of
    what is free at creation)."""

    def __init__(self, p: Params | None = None, slots: int = 4, device: int = 0,
                 hbm_budget: int | None = None):
        self.params = p if p is not None else _params()
        self._h = C.c_void_p()
        st = lib().ar_server_create(C.byref(self.params), int(device), int(slots),
                                    int(hbm_budget or 0), C.byref(self._h))
        if st != _lib.AR_OK:
            raise ServerError(st, "server creation failed (no device, or bad slot count)")
        self._pending = {}

    def close(self):
        if self._h:
            lib().ar_server_destroy(self._h)
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

    def _check(self, st):
        if st == _lib.AR_OK:
            return
        msg = lib().ar_server_last_error(self._h).decode(errors="replace")
        if st == _lib.AR_ERR_INVALID_ARGUMENT:
            raise _lib.ArgumentError(msg)
        if st == _lib.AR_ERR_DIMENSION_MISMATCH:
            raise _lib.DimensionMismatch(msg)
        raise ServerError(st, msg)

    def submit(self, reference, secondary, reference_valid=None, secondary_valid=None,
               grid=None) -> int:
        """Queue one pair; returns its ticket. The inputs are copied before this returns."""
        ref = _fortran(reference)
        sec = _fortran(secondary, ref.dtype)
        if ref.shape != sec.shape:
            raise _lib.DimensionMismatch(f"reference is {ref.shape} but secondary is {sec.shape}")
        if ref.dtype not in _DTYPES:
            raise _lib.ArgumentError(f"unsupported image element type {ref.dtype}")
        rv = None if reference_valid is None else _fortran(reference_valid, np.uint8)
        sv = None if secondary_valid is None else _fortran(secondary_valid, np.uint8)
        g = _build_grid(ref.shape, self.params) if grid is None else grid
        if len(np.shape(g.x)) != 2:
            raise _lib.ArgumentError("the served search needs a gridded (2-D) point set")
        s, keep = _points_struct(g)
        nr, nc = np.shape(g.x)
        t = C.c_int64()
        ny, nx = ref.shape
        self._check(lib().ar_server_submit(self._h, _ptr(ref), _ptr(sec), _ptr(rv), _ptr(sv),
                                           ny, nx, _DTYPES[ref.dtype], C.byref(s), nr, nc,
                                           C.byref(t)))
        self._pending[t.value] = (nr, nc)
        return t.value

    def result(self, ticket: int):
        """Block until the pair is done: (MultichipResult, timing) where timing holds
        latency_ms (submission to completion), queued_ms, service_ms and slot."""
        nr, nc = self._pending.pop(ticket)
        n = nr * nc
        dx = np.empty(n, np.float32)
        dy = np.empty(n, np.float32)
        cc = np.empty(n, np.float32)
        cs = np.empty(n, np.uint16)
        it = np.empty(n, np.uint8)
        tm = (C.c_double * 4)()
        self._check(lib().ar_server_wait(self._h, int(ticket), _ptr(dx), _ptr(dy), _ptr(cc),
                                         _ptr(cs), _ptr(it), tm))
        r = lambda a: a.reshape((nr, nc), order="F")
        timing = {"latency_ms": tm[0], "queued_ms": tm[1], "service_ms": tm[2],
                  "slot": int(tm[3])}
        return MultichipResult(r(dx), r(dy), r(cc), r(cs), r(it).astype(bool)), timing

    def stats(self) -> dict:
        v = (C.c_int64 * 8)()
        self._check(lib().ar_server_stats(self._h, v))
        keys = ("completed", "queued", "running", "peak_running", "bytes_held",
                "peak_bytes_held", "budget", "slots")
        return dict(zip(keys, [int(x) for x in v]))
