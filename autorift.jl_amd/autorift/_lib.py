"""ctypes binding of libautorift_hip.so (the C ABI declared in include/autorift_hip.h).

The library is the product: there is no CPU fallback. Loading fails loudly when the shared
library has not been built, and calls fail with the library's own status and message.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libautorift_hip.so")

AR_OK, AR_ERR_INVALID_ARGUMENT, AR_ERR_DIMENSION_MISMATCH = 0, 1, 2
AR_ERR_OUT_OF_MEMORY, AR_ERR_DEVICE, AR_ERR_UNSUPPORTED, AR_ERR_NO_DEVICE = 3, 4, 5, 6
AR_U8, AR_F32, AR_F64, AR_I16, AR_C64 = 0, 1, 2, 3, 4
AR_ZNCC, AR_NCC, AR_COHERENCE = 0, 1, 2
AR_PREC_EXACT, AR_PREC_BF16, AR_PREC_FP16 = 0, 1, 2
AR_BATCH_GRAPH, AR_BATCH_SPLIT_SET = 1, 2


class AutoRIFTError(RuntimeError):
    """A device-side failure (allocation, runtime, unsupported geometry)."""

    def __init__(self, status, message):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class ArgumentError(ValueError):
    """Invalid argument, as Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Incompatible array sizes, as Julia's DimensionMismatch."""


class NoDeviceError(AutoRIFTError):
    pass


class Params(C.Structure):
    """ar_params: the flattened Params of src/types.jl:781-829."""
    _fields_ = [
        ("preprocess", C.c_int32), ("filter_width", C.c_int32), ("quantize", C.c_int32),
        ("measure", C.c_int32), ("upsampling", C.c_int64), ("outliers", C.c_int32),
        ("_pad0", C.c_int32), ("gardner_window", C.c_int64), ("gardner_iterations", C.c_int64),
        ("min_agree_fraction", C.c_double), ("agree_tolerance", C.c_double),
        ("mad_scale", C.c_double), ("chip_size_base", C.c_int64), ("chip_size_min", C.c_int64),
        ("chip_size_max", C.c_int64), ("chip_aspect", C.c_double), ("grid_spacing", C.c_int64),
        ("search_radius_x", C.c_int64), ("search_radius_y", C.c_int64),
        ("min_search_radius", C.c_int64), ("coarse_stride", C.c_int64),
        ("coarse_buffer", C.c_int64), ("min_coarse_valid_fraction", C.c_double),
        ("dx_prior", C.c_double), ("dy_prior", C.c_double), ("fill_window", C.c_int64),
        ("precision", C.c_int32), ("rotation_count", C.c_int32),
        ("rotation_angles", C.c_double * 8), ("deramp_axis", C.c_int32),
        ("level_measure_count", C.c_int32), ("level_measures", C.c_int32 * 8),
        ("min_std", C.c_double),
    ]


class Points(C.Structure):
    """ar_points: PointSet fields (src/points.jl:50-74) as pointers."""
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "radius_x", "radius_y", "dx_prior",
                                          "dy_prior", "chip_size_x", "chip_size_y")]


# (name, restype, argtypes) for every symbol declared in include/autorift_hip.h
SIGNATURES = [
    ("ar_params_default", None, [C.POINTER(Params)]),
    ("ar_context_create", C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    ("ar_context_destroy", None, [C.c_void_p]),
    ("ar_set_images", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64, C.c_int64, C.c_int32]),
    ("ar_set_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_int64, C.c_int64, C.c_int32]),
    ("ar_get_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.POINTER(C.c_int32)]),
    ("ar_quantizer_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("ar_sequential_sum", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                    C.POINTER(C.c_double), C.c_void_p]),
    ("ar_track", C.c_int, [C.c_void_p, C.POINTER(Points), C.c_int64, C.c_int64, C.c_int32,
                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("ar_track_device", C.c_int, [C.c_void_p, C.POINTER(Points), C.c_int64, C.c_int64,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("ar_run_multichip", C.c_int, [C.c_void_p, C.POINTER(Points), C.c_int64, C.c_int64,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    ("ar_correlate_point", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int64,
                                     C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_float)]),
    ("ar_correlate_point_noalloc", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
      C.c_int64, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    ("ar_last_kernel", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("ar_last_timing", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    ("ar_last_error", C.c_char_p, [C.c_void_p]),
    ("ar_version", C.c_char_p, []),
    ("ar_batch_create", C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    ("ar_batch_destroy", None, [C.c_void_p]),
    ("ar_batch_add_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_int32)]),
    ("ar_batch_track", C.c_int, [C.c_void_p, C.POINTER(Points), C.c_void_p, C.c_int64,
                                 C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    ("ar_batch_track_device", C.c_int, [C.c_void_p, C.POINTER(Points), C.c_void_p, C.c_int64,
                                        C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int32]),
    ("ar_batch_info", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("ar_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("ar_batch_set_exact", C.c_int, [C.c_void_p, C.c_int32]),
    ("ar_tune_cache_path", C.c_int64, [C.c_char_p, C.c_char_p, C.c_int64]),
    ("ar_tune_cache_lookup", C.c_int32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32)]),
    ("ar_tune_cache_record", C.c_int32, [C.c_char_p, C.c_char_p, C.c_int32]),
    ("ar_server_create", C.c_int, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int64,
                                   C.POINTER(C.c_void_p)]),
    ("ar_server_destroy", None, [C.c_void_p]),
    ("ar_server_submit", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.c_int64, C.c_int32, C.POINTER(Points), C.c_int64,
                                   C.c_int64, C.POINTER(C.c_int64)]),
    ("ar_server_wait", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    ("ar_server_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("ar_server_last_error", C.c_char_p, [C.c_void_p]),
]

_lib = None


def lib():
    """The loaded library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(or `make -C autorift.jl_amd`)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status == AR_OK:
        return
    msg = lib().ar_last_error(ctx).decode(errors="replace")
    if status == AR_ERR_INVALID_ARGUMENT:
        raise ArgumentError(msg)
    if status == AR_ERR_DIMENSION_MISMATCH:
        raise DimensionMismatch(msg)
    if status == AR_ERR_NO_DEVICE:
        raise NoDeviceError(status, msg)
    raise AutoRIFTError(status, msg)


def default_params() -> Params:
    p = Params()
    lib().ar_params_default(C.byref(p))
    return p
