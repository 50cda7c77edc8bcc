"""MI355X-native per-timestep rigid-body contact pipeline (grid-SDF contacts,
primal-dual friction Newton, block-Jacobi PCG, merit line search) behind a C ABI.

The compute path is libpdc.so (hand-written HIP for gfx950); this package holds
its ctypes binding, the parameter structures and the seeded synthetic scenes.
"""
from .params import (NON_SMOOTH_CONTACT, ONLY_NORMAL, PCG_FP32, PCG_FP64, SHAPE_BOX, SHAPE_ELLIPSOID, SHAPE_SPHERE,
                     PdcParams, PdcSolveStats, PdcTimings, default_params)
from .pdc import (ConfigError, Context, DynamicSystem, PdcError, PrimalDualForceSolver, device_count,
                  load_library, LIB_PATH)

__all__ = [
    "NON_SMOOTH_CONTACT", "ONLY_NORMAL", "PCG_FP32", "PCG_FP64", "SHAPE_BOX", "SHAPE_ELLIPSOID", "SHAPE_SPHERE",
    "PdcParams", "PdcSolveStats", "PdcTimings", "default_params", "ConfigError", "Context",
    "DynamicSystem", "PdcError", "PrimalDualForceSolver", "device_count", "load_library", "LIB_PATH",
]
