"""MI355X-native dense ZNCC tracker with AutoRIFT.jl's interface.

Python host over the C ABI of libautorift_hip.so (include/autorift_hip.h). Names follow the
reference (src/api.jl, src/points.jl, src/track.jl, src/multichip.jl); Julia's mutating
`reinit!` / `autorift!` are spelled `reinit_` / `autorift_`.
"""
from ._lib import (AR_COHERENCE, AR_NCC, AR_ZNCC, ArgumentError, AutoRIFTError, DimensionMismatch,
                   NoDeviceError, Params, default_params, lib)
from .api import (Cache, Highpass, ImagePair, RotationSearch, Wallis, angles, autorift, autorift_,
                  autorift_with_grid, chip_measures, chip_sizes, init, measure_at, params,
                  reinit_, track)
from .core import Context, DisplacementField, MultichipResult
from .points import (PointSet, chip_bounds, gridpoints, gridpoints_xy, pointset, sanitize_,
                     search_bounds)
from .batch import Batch
from .geo import autorift_geo, grid_coordinates, pixel_size, to_velocity, years
from .serve import Server, ServerError

__all__ = [
    "AR_COHERENCE", "AR_NCC", "AR_ZNCC", "ArgumentError", "AutoRIFTError", "DimensionMismatch", "NoDeviceError",
    "Params", "default_params", "lib", "Cache", "ImagePair", "autorift", "autorift_",
    "autorift_with_grid", "chip_sizes", "init", "params", "reinit_", "track", "Context",
    "DisplacementField", "MultichipResult", "PointSet", "chip_bounds", "gridpoints",
    "gridpoints_xy", "pointset", "sanitize_", "search_bounds", "Server", "ServerError", "Batch", "RotationSearch", "angles", "autorift_geo", "grid_coordinates",
    "pixel_size", "to_velocity", "years", "chip_measures", "measure_at", "Highpass", "Wallis",
]
