"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference's
own .ini reader built into oracle/_ref/libref_config.so.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from euler2d_kokkos_amd._capi import E2DParams

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
REF_CONFIG_LIB = ORACLE_DIR / "_ref" / "libref_config.so"

_D4 = C.c_double * 4
_P = C.POINTER(E2DParams)


def _build_if_missing() -> None:
    if not ORACLE_LIB.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR), "oracle"], check=True,
                       capture_output=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    _build_if_missing()
    L = C.CDLL(str(ORACLE_LIB))
    protos = {
        "ora_params_size": (C.c_int, []),
        "ora_config_open": (C.c_void_p, [C.c_char_p, C.POINTER(C.c_int)]),
        "ora_config_close": (None, [C.c_void_p]),
        "ora_config_get_string": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                            C.c_char_p, C.c_int]),
        "ora_config_get_integer": (C.c_long, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_long]),
        "ora_config_get_float": (C.c_float, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_float]),
        "ora_config_get_bool": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]),
        "ora_params_defaults": (None, [_P]),
        "ora_params_setup": (None, [C.c_void_p, _P]),
        "ora_params_init": (None, [_P]),
        "ora_compute_primitives": (None, [_P, _D4, C.POINTER(C.c_double), _D4]),
        "ora_slope_unsplit_hydro_2d": (None, [_P, _D4, _D4, _D4, _D4, _D4, _D4, _D4]),
        "ora_trace_unsplit_2d_along_dir": (None, [_P, _D4, _D4, _D4, C.c_double, C.c_double,
                                                  C.c_int, _D4]),
        "ora_riemann_hllc": (None, [_P, _D4, _D4, _D4]),
        "ora_riemann_hll": (None, [_P, _D4, _D4, _D4]),
        "ora_riemann_approx": (None, [_P, _D4, _D4, _D4, _D4]),
        "ora_create": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
        "ora_destroy": (None, [C.c_void_p]),
        "ora_set_threads": (None, [C.c_int]),
        "ora_get_threads": (C.c_int, []),
        "ora_init_problem": (None, [C.c_void_p]),
        "ora_compute_dt": (C.c_double, [C.c_void_p, C.c_int]),
        "ora_make_boundaries": (None, [C.c_void_p, C.c_int]),
        "ora_convert_to_primitives": (None, [C.c_void_p, C.c_int]),
        "ora_compute_and_store_fluxes": (None, [C.c_void_p, C.c_double, C.c_double]),
        "ora_update": (None, [C.c_void_p, C.c_int]),
        "ora_godunov_unsplit": (None, [C.c_void_p, C.c_int, C.c_double]),
        "ora_run_loop": (None, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]),
        "ora_get_array": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
        "ora_set_array": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    }
    for name, (res, args) in protos.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    assert L.ora_params_size() == C.sizeof(E2DParams)
    _lib = L
    return L


def d4(x) -> C.Array:
    return _D4(*[float(v) for v in x])


def params_from_ini(path) -> E2DParams:
    L = lib()
    err = C.c_int(0)
    h = L.ora_config_open(str(path).encode(), C.byref(err))
    p = E2DParams()
    L.ora_params_setup(h, C.byref(p))
    L.ora_config_close(h)
    return p


def params_from_text(tmp_path: Path, text: str, name: str = "deck.ini") -> E2DParams:
    f = Path(tmp_path) / name
    f.write_text(text)
    return params_from_ini(f)


def default_params(**over) -> E2DParams:
    """HydroParams after setup() on an empty deck, then field overrides + init()."""
    L = lib()
    h = L.ora_config_open(b"/nonexistent/deck.ini", None)
    p = E2DParams()
    L.ora_params_setup(h, C.byref(p))
    L.ora_config_close(h)
    for k, v in over.items():
        setattr(p, k, v)
    L.ora_params_init(C.byref(p))
    return p


# ------------------------------------------------------------- pointwise
def prim(p: E2DParams, u):
    q = _D4()
    c = C.c_double()
    lib().ora_compute_primitives(C.byref(p), d4(u), C.byref(c), q)
    return np.array(q[:]), c.value


def slopes(p, q, qpx, qmx, qpy, qmy):
    dx, dy = _D4(), _D4()
    lib().ora_slope_unsplit_hydro_2d(C.byref(p), d4(q), d4(qpx), d4(qmx), d4(qpy), d4(qmy),
                                     dx, dy)
    return np.array(dx[:]), np.array(dy[:])


def trace(p, q, dqx, dqy, dtdx, dtdy, face):
    out = _D4()
    lib().ora_trace_unsplit_2d_along_dir(C.byref(p), d4(q), d4(dqx), d4(dqy), dtdx, dtdy,
                                         face, out)
    return np.array(out[:])


def hllc(p, ql, qr):
    f = _D4()
    lib().ora_riemann_hllc(C.byref(p), d4(ql), d4(qr), f)
    return np.array(f[:])


def hll(p, ql, qr):
    f = _D4()
    lib().ora_riemann_hll(C.byref(p), d4(ql), d4(qr), f)
    return np.array(f[:])


def approx(p, ql, qr):
    g, f = _D4(), _D4()
    lib().ora_riemann_approx(C.byref(p), d4(ql), d4(qr), g, f)
    return np.array(g[:]), np.array(f[:])


# ------------------------------------------------------------------- run
class OracleRun:
    """HydroRun (implementation version 0) on the CPU; arrays are [i][j][var]."""

    def __init__(self, p: E2DParams, init: bool = True):
        self.p = p.copy()
        self._h = C.c_void_p()
        rc = lib().ora_create(C.byref(self.p), C.byref(self._h))
        if rc != 0:
            raise ValueError(f"ora_create failed ({rc})")
        self.t = 0.0
        self.nStep = 0
        if init:
            lib().ora_init_problem(self._h)

    def close(self):
        if self._h:
            lib().ora_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self.p.isize, self.p.jsize, 4)

    def get(self, which: int) -> np.ndarray:
        a = np.empty(self.shape, dtype=np.float64)
        lib().ora_get_array(self._h, which, a.ctypes.data)
        return a

    def set(self, which: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        lib().ora_set_array(self._h, which, a.ctypes.data)

    def compute_dt(self, useU: int) -> float:
        return lib().ora_compute_dt(self._h, useU)

    def make_boundaries(self, which: int) -> None:
        lib().ora_make_boundaries(self._h, which)

    def godunov_unsplit(self, nStep: int, dt: float) -> None:
        lib().ora_godunov_unsplit(self._h, nStep, dt)

    def run(self, max_steps: int = 0):
        t = C.c_double(self.t)
        n = C.c_int(self.nStep)
        lib().ora_run_loop(self._h, C.byref(t), C.byref(n), max_steps)
        self.t, self.nStep = t.value, n.value
        return self.t, self.nStep

    def current(self) -> np.ndarray:
        return self.get(0 if self.nStep % 2 == 0 else 1)


def interior(a: np.ndarray) -> np.ndarray:
    """[i][j][var] with ghosts -> interior block."""
    return a[2:-2, 2:-2, :]


def rel_l1(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-variable relative L1 = sum|a-b| / sum|b| over the given cells."""
    a = a.reshape(-1, 4)
    b = b.reshape(-1, 4)
    return np.abs(a - b).sum(axis=0) / np.maximum(np.abs(b).sum(axis=0), 1e-300)


# ----------------------------------------------- reference .ini reader
_ref = None


def ref_config_lib():
    """The reference's own ConfigMap/INIReader/ini.cpp, or None if not built."""
    global _ref
    if _ref is not None:
        return _ref
    if not REF_CONFIG_LIB.exists():
        return None
    L = C.CDLL(str(REF_CONFIG_LIB))
    L.refcfg_open.restype = C.c_void_p
    L.refcfg_open.argtypes = [C.c_char_p]
    L.refcfg_parse_error.restype = C.c_int
    L.refcfg_parse_error.argtypes = [C.c_void_p]
    L.refcfg_close.restype = None
    L.refcfg_close.argtypes = [C.c_void_p]
    L.refcfg_get_string.restype = C.c_int
    L.refcfg_get_string.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                    C.c_int]
    L.refcfg_get_integer.restype = C.c_long
    L.refcfg_get_integer.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_long]
    L.refcfg_get_float.restype = C.c_float
    L.refcfg_get_float.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_float]
    L.refcfg_get_bool.restype = C.c_int
    L.refcfg_get_bool.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
    _ref = L
    return L


def _output_protos() -> None:
    L = lib()
    if getattr(L, "_output_ready", False):
        return
    L.ora_save_vtk.restype = C.c_int
    L.ora_save_vtk.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    L.ora_radial_profile.restype = C.c_int
    L.ora_radial_profile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    L.ora_write_npy.restype = C.c_int
    L.ora_write_npy.argtypes = [C.c_char_p, C.c_void_p, C.c_int64]
    L._output_ready = True


def save_vtk(run: "OracleRun", which: int, output_dir, prefix: str, step: int,
             float32: bool = False) -> Path:
    _output_protos()
    rc = lib().ora_save_vtk(run._h, which, int(float32), str(output_dir).encode(),
                            prefix.encode(), step)
    assert rc == 0, rc
    return Path(output_dir) / f"{prefix}_{step:07d}.vti"


def radial_profile(run: "OracleRun", which: int, nbins: int):
    _output_protos()
    d = np.empty(nbins)
    m = np.empty(nbins)
    c = np.empty(nbins, dtype=np.int64)
    rc = lib().ora_radial_profile(run._h, which, nbins, d.ctypes.data, m.ctypes.data,
                                  c.ctypes.data)
    assert rc == 0, rc
    return d, m, c


def write_npy(path, data: np.ndarray) -> None:
    _output_protos()
    a = np.ascontiguousarray(data, dtype=np.float64).ravel()
    assert lib().ora_write_npy(str(path).encode(), a.ctypes.data, a.size) == 0
