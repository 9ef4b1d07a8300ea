"""MI355X-native unsplit MUSCL-Hancock Godunov solver (euler2d_kokkos drop-in).

The compute path is libeuler2d_hip.so (hand-written HIP for gfx950) behind the
C ABI in include/euler2d_hip.h; this package is its Python host mirror.
"""
from ._capi import (E2D_BATCH_LAUNCH_PERSISTENT, E2D_BATCH_LAUNCH_STEPS, E2D_LAYOUT_INTERIOR,
                    E2D_LAYOUT_LEFT, E2D_LAYOUT_RIGHT,
                    E2D_PRECISION_FP32, E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64,
                    E2D_SCHEDULE_ATOMIC, E2D_SCHEDULE_DIRECTIONAL, E2D_SCHEDULE_FUSED,
                    E2D_SCHEDULE_PASSES, E2D_U, E2D_U2, E2DError, E2DParams, load_library)
from .hydro import (ID, IP, IU, IV, PROBLEMS, RIEMANN, ConfigMap, HydroParams, HydroRun,
                    device_info, write_npy, write_vti)

__all__ = [
    "E2D_BATCH_LAUNCH_PERSISTENT", "E2D_BATCH_LAUNCH_STEPS", "E2D_LAYOUT_INTERIOR", "E2D_LAYOUT_LEFT", "E2D_LAYOUT_RIGHT", "E2D_PRECISION_FP32",
    "E2D_PRECISION_FP32_SPLIT",
    "E2D_PRECISION_FP64", "E2D_SCHEDULE_ATOMIC", "E2D_SCHEDULE_DIRECTIONAL", "E2D_SCHEDULE_FUSED",
    "E2D_SCHEDULE_PASSES", "E2D_U", "E2D_U2",
    "E2DError", "E2DParams", "load_library", "ID", "IP", "IU", "IV", "PROBLEMS", "RIEMANN",
    "ConfigMap", "HydroParams", "HydroRun", "device_info", "write_npy", "write_vti",
]
