"""ctypes binding of libeuler2d_hip.so (include/euler2d_hip.h).

The library is built in-tree by ``python -m euler2d_kokkos_amd.build``.  There
is no fallback: if the shared library is missing or fails to load, every call
raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
# E2D_LIBRARY names another build of the library (a variant built with other
# compile-time options); by default the in-tree build
LIB_PATH = Path(os.environ.get("E2D_LIBRARY") or PKG / "libeuler2d_hip.so")

E2D_OK = 0
E2D_ERR_ARG = -1
E2D_ERR_IO = -2
E2D_ERR_HIP = -3
E2D_ERR_ALLOC = -4
E2D_ERR_UNSUPPORTED = -5

E2D_U = 0
E2D_U2 = 1

E2D_BATCH_LAUNCH_STEPS = 0
E2D_BATCH_LAUNCH_PERSISTENT = 1
E2D_LAYOUT_RIGHT = 0
E2D_LAYOUT_LEFT = 1
E2D_LAYOUT_INTERIOR = 2

E2D_PRECISION_FP64 = 0
E2D_PRECISION_FP32_SPLIT = 1
E2D_PRECISION_FP32 = 2

E2D_SCHEDULE_FUSED = 0
E2D_SCHEDULE_PASSES = 1
E2D_SCHEDULE_DIRECTIONAL = 2
E2D_SCHEDULE_ATOMIC = 3

_I = C.c_int32
_D = C.c_double

# Field order and types of struct e2d_params (include/euler2d_hip.h).
PARAM_FIELDS = [
    ("nStepmax", _I), ("nOutput", _I), ("enableOutput", _I), ("ioVTK", _I), ("ioHDF5", _I),
    ("ghostWidth", _I), ("tEnd", _D),
    ("nx", _I), ("ny", _I), ("imin", _I), ("imax", _I), ("jmin", _I), ("jmax", _I),
    ("isize", _I), ("jsize", _I),
    ("boundary_type_xmin", _I), ("boundary_type_xmax", _I),
    ("boundary_type_ymin", _I), ("boundary_type_ymax", _I),
    ("xmin", _D), ("xmax", _D), ("ymin", _D), ("ymax", _D), ("dx", _D), ("dy", _D),
    ("gamma0", _D), ("gamma6", _D), ("cfl", _D), ("slope_type", _D), ("smallr", _D),
    ("smallc", _D), ("smallp", _D), ("smallpp", _D),
    ("niter_riemann", _I), ("riemannSolverType", _I), ("problemType", _I),
    ("implementationVersion", _I), ("honor_riemann", _I), ("blast_nbins", _I),
    ("blast_radius", _D), ("blast_center_x", _D), ("blast_center_y", _D),
    ("blast_density_in", _D), ("blast_density_out", _D),
    ("blast_pressure_in", _D), ("blast_pressure_out", _D), ("blast_total_energy_inside", _D),
    ("sb_bubble_radius", _D), ("sb_bubble_center_x", _D), ("sb_bubble_center_y", _D),
    ("sb_bubble_density", _D), ("sb_bubble_pressure", _D),
    ("sb_preshock_density", _D), ("sb_preshock_pressure", _D),
    ("sb_postshock_density", _D), ("sb_postshock_pressure", _D), ("sb_postshock_velocity", _D),
    ("sb_shock_loc", _D),
    ("implode_pressure_in", _D),
    ("kin_frac", _D),
]


class E2DParams(C.Structure):
    _fields_ = PARAM_FIELDS

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in PARAM_FIELDS}

    def copy(self) -> "E2DParams":
        other = E2DParams()
        C.pointer(other)[0] = self
        return other


class E2DError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"euler2d error {code}: {message}")
        self.code = code


_PROTOTYPES = {
    "e2d_last_error": (C.c_char_p, []),
    "e2d_version": (C.c_int, []),
    "e2d_params_size": (C.c_int, []),
    "e2d_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "e2d_config_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    "e2d_config_close": (C.c_int, [C.c_void_p]),
    "e2d_config_get_string": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                        C.c_char_p, C.c_int]),
    "e2d_config_get_integer": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_long,
                                         C.POINTER(C.c_long)]),
    "e2d_config_get_float": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_float,
                                       C.POINTER(C.c_float)]),
    "e2d_config_get_bool": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int,
                                      C.POINTER(C.c_int)]),
    "e2d_params_setup": (C.c_int, [C.c_void_p, C.POINTER(E2DParams)]),
    "e2d_params_from_ini": (C.c_int, [C.c_char_p, C.POINTER(E2DParams), C.POINTER(C.c_int)]),
    "e2d_params_init": (C.c_int, [C.POINTER(E2DParams)]),
    "e2d_params_print": (C.c_int, [C.POINTER(E2DParams)]),
    "e2d_params_validate": (C.c_int, [C.POINTER(E2DParams)]),
    "e2d_create": (C.c_int, [C.POINTER(E2DParams), C.c_int, C.POINTER(C.c_void_p)]),
    "e2d_destroy": (C.c_int, [C.c_void_p]),
    "e2d_init_problem": (C.c_int, [C.c_void_p]),
    "e2d_compute_dt": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "e2d_make_boundaries": (C.c_int, [C.c_void_p, C.c_int]),
    "e2d_godunov_unsplit": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "e2d_run_loop": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "e2d_run_loop_noalloc": (C.c_int, [C.POINTER(E2DParams), C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                                       C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p]),
    "e2d_copy_state": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "e2d_upload_state": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "e2d_copy_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "e2d_conserved_sums": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "e2d_timers": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "e2d_kinetic_cells": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]),
    "e2d_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "e2d_save_vtk": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    "e2d_write_vti": (C.c_int, [C.POINTER(E2DParams), C.c_void_p, C.c_int, C.c_char_p,
                                C.c_char_p, C.c_int]),
    "e2d_radial_profile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "e2d_write_npy": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64]),
    "e2d_set_schedule": (C.c_int, [C.c_void_p, C.c_int]),
    "e2d_set_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "e2d_get_tile": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "e2d_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "e2d_synchronize": (C.c_int, [C.c_void_p]),
    "e2d_time_loop": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "e2d_batch_create": (C.c_int, [C.POINTER(E2DParams), C.c_int, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    "e2d_batch_destroy": (C.c_int, [C.c_void_p]),
    "e2d_batch_init": (C.c_int, [C.c_void_p]),
    "e2d_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "e2d_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "e2d_batch_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "e2d_batch_queue_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "e2d_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_int)]),
    "e2d_batch_status": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]),
    "e2d_batch_copy_state": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "e2d_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
    "e2d_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_double)]),
}

EXPORTED = tuple(_PROTOTYPES)

_lib = None


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Loads (once) and returns the HIP solver library; raises when absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m euler2d_kokkos_amd.build` "
            "(the HIP path has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.e2d_params_size() != C.sizeof(E2DParams):
        raise RuntimeError("e2d_params layout mismatch between the library and the binding")
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> int:
    """Raises E2DError for a negative status code."""
    if rc < 0:
        msg = load_library().e2d_last_error()
        raise E2DError(rc, msg.decode() if msg else "")
    return rc
