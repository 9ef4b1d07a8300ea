"""Python mirror of the reference's solver interface, over the C ABI.

Names, argument meaning and error behaviour follow pkestene/euler2d_kokkos:
  ConfigMap    config/ConfigMap.h:25-47, config/inih/INIReader.h:47-67
  HydroParams  src/HydroParams.h:155-265 (setup/init/print, HydroParams.cpp:42-224)
  HydroRun     src/HydroRun.h:44-134 (compute_dt :227, godunov_unsplit :260,
               make_boundaries :390, members U / U2)
Arrays cross the boundary as numpy float64 arrays; the default layout is the
reference's OpenMP View layout [i][j][var] with ghost cells (LayoutRight).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _capi
from ._capi import E2DParams, check, load_library

ID, IP, IU, IV = 0, 1, 2, 3
PROBLEMS = {"implode": 0, "blast": 1, "four_quadrant": 2, "discontinuity": 3,
            "shocked_bubble": 4}
RIEMANN = {"approx": 0, "hll": 1, "hllc": 2}


class ConfigMap:
    """Parsed .ini file with ConfigMap / INIReader getters."""

    def __init__(self, filename: str | Path):
        lib = load_library()
        self._h = C.c_void_p()
        err = C.c_int(0)
        rc = lib.e2d_config_open(str(filename).encode(), C.byref(self._h), C.byref(err))
        self._error = err.value
        if rc not in (_capi.E2D_OK, _capi.E2D_ERR_IO):
            check(rc)

    def ParseError(self) -> int:
        return self._error

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load_library().e2d_config_close(h)
            self._h = None

    def getString(self, section: str, name: str, default_value: str) -> str:
        lib = load_library()
        n = check(lib.e2d_config_get_string(self._h, section.encode(), name.encode(),
                                            default_value.encode(), None, 0))
        buf = C.create_string_buffer(n + 1)
        check(lib.e2d_config_get_string(self._h, section.encode(), name.encode(),
                                        default_value.encode(), buf, n + 1))
        return buf.value.decode()

    def getInteger(self, section: str, name: str, default_value: int) -> int:
        out = C.c_long(0)
        check(load_library().e2d_config_get_integer(self._h, section.encode(), name.encode(),
                                                    default_value, C.byref(out)))
        return out.value

    def getFloat(self, section: str, name: str, default_value: float) -> float:
        out = C.c_float(0)
        check(load_library().e2d_config_get_float(self._h, section.encode(), name.encode(),
                                                  default_value, C.byref(out)))
        return out.value

    def getBool(self, section: str, name: str, default_value: bool) -> bool:
        out = C.c_int(0)
        check(load_library().e2d_config_get_bool(self._h, section.encode(), name.encode(),
                                                 int(bool(default_value)), C.byref(out)))
        return bool(out.value)


class HydroParams:
    """HydroParams: every field of the reference struct as an attribute."""

    def __init__(self, raw: E2DParams | None = None):
        object.__setattr__(self, "_raw", raw if raw is not None else E2DParams())

    @classmethod
    def from_ini(cls, filename: str | Path) -> "HydroParams":
        p = cls()
        p.setup(ConfigMap(filename))
        return p

    def setup(self, config: ConfigMap) -> None:
        check(load_library().e2d_params_setup(config._h, C.byref(self._raw)))

    def init(self) -> None:
        check(load_library().e2d_params_init(C.byref(self._raw)))

    def print(self) -> None:
        check(load_library().e2d_params_print(C.byref(self._raw)))

    def validate(self) -> None:
        check(load_library().e2d_params_validate(C.byref(self._raw)))

    @property
    def raw(self) -> E2DParams:
        return self._raw

    def __getattr__(self, name):
        return getattr(object.__getattribute__(self, "_raw"), name)

    def __setattr__(self, name, value):
        setattr(self._raw, name, value)

    def as_dict(self) -> dict:
        return self._raw.as_dict()


class HydroRun:
    """Solver state on the GPU: U and U2 with ghost cells, one HIP stream.

    Unlike the reference constructor, initialisation of the problem is the
    separate call init_problem(); HydroRun(params) performs both unless
    ``initialise=False``.
    """

    def __init__(self, params: HydroParams, precision: int = _capi.E2D_PRECISION_FP64,
                 schedule: int | None = None, initialise: bool = True):
        """schedule None keeps the one implementationVersion selects."""
        lib = load_library()
        self.params = params
        self._h = C.c_void_p()
        check(lib.e2d_create(C.byref(params.raw), precision, C.byref(self._h)))
        if schedule is not None:
            check(lib.e2d_set_schedule(self._h, schedule))
        self.t = 0.0
        self.nStep = 0
        if initialise:
            self.init_problem()

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().e2d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference API
    def init_problem(self) -> None:
        check(load_library().e2d_init_problem(self._h))
        self.t, self.nStep = 0.0, 0

    def compute_dt(self, useU: int) -> float:
        dt = C.c_double(0.0)
        check(load_library().e2d_compute_dt(self._h, useU, C.byref(dt)))
        return dt.value

    def make_boundaries(self, which: int) -> None:
        check(load_library().e2d_make_boundaries(self._h, which))

    def godunov_unsplit(self, nStep: int, dt: float) -> None:
        check(load_library().e2d_godunov_unsplit(self._h, nStep, dt))

    def set_schedule(self, schedule: int) -> None:
        check(load_library().e2d_set_schedule(self._h, schedule))

    def get_schedule(self) -> int:
        out = C.c_int(-1)
        check(load_library().e2d_get_schedule(self._h, C.byref(out)))
        return out.value

    # -- main.cpp time loop with device-resident dt
    def run(self, max_steps: int = 0) -> tuple[float, int]:
        t = C.c_double(self.t)
        n = C.c_int(self.nStep)
        check(load_library().e2d_run_loop(self._h, max_steps, C.byref(t), C.byref(n)))
        self.t, self.nStep = t.value, n.value
        return self.t, self.nStep

    def time_loop(self, nsteps: int) -> tuple[float, int]:
        """Runs nsteps loop iterations; returns (device milliseconds, steps run)."""
        ms = C.c_double(0.0)
        done = C.c_int(0)
        t = C.c_double(self.t)
        n = C.c_int(self.nStep)
        check(load_library().e2d_time_loop(self._h, nsteps, C.byref(ms), C.byref(done),
                                           C.byref(t), C.byref(n)))
        self.t, self.nStep = t.value, n.value
        return ms.value, done.value

    def set_tile(self, seg_rows: int) -> None:
        check(load_library().e2d_set_tile(self._h, seg_rows))

    def get_tile(self) -> int:
        out = C.c_int(0)
        check(load_library().e2d_get_tile(self._h, C.byref(out)))
        return out.value

    def autotune(self, candidates=None, steps: int = 10):
        """Times each candidate tile on the current state (restored afterwards)
        and keeps the fastest; returns (seg_rows, ms per step).  candidates
        None: a fixed set plus the tiles at which the wavefronts fit k per
        SIMD (k = 1..4)."""
        n = 0 if candidates is None else len(candidates)
        cand = (C.c_int * max(n, 1))(*(candidates or ()))
        best = C.c_int(0)
        ms = C.c_double(0.0)
        check(load_library().e2d_autotune(self._h, cand if n else None, n, steps, C.byref(best),
                                          C.byref(ms)))
        return best.value, ms.value

    def synchronize(self) -> None:
        check(load_library().e2d_synchronize(self._h))

    # -- data access
    def _shape(self, layout: int) -> tuple[int, ...]:
        p = self.params
        if layout == _capi.E2D_LAYOUT_RIGHT:
            return (p.isize, p.jsize, 4)
        if layout == _capi.E2D_LAYOUT_LEFT:
            return (4, p.jsize, p.isize)
        return (4, p.ny, p.nx)

    def get_state(self, which: int, layout: int = _capi.E2D_LAYOUT_RIGHT) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=np.float64)
        check(load_library().e2d_copy_state(self._h, which, out.ctypes.data, layout))
        return out

    def set_state(self, which: int, array: np.ndarray,
                  layout: int = _capi.E2D_LAYOUT_RIGHT) -> None:
        a = np.ascontiguousarray(array, dtype=np.float64)
        if a.shape != self._shape(layout):
            raise ValueError(f"expected shape {self._shape(layout)}, got {a.shape}")
        check(load_library().e2d_upload_state(self._h, which, a.ctypes.data, layout))

    @property
    def U(self) -> np.ndarray:
        return self.get_state(_capi.E2D_U)

    @property
    def U2(self) -> np.ndarray:
        return self.get_state(_capi.E2D_U2)

    def current(self) -> int:
        """Selector of the state after nStep steps (main.cpp:155-158)."""
        return _capi.E2D_U if self.nStep % 2 == 0 else _capi.E2D_U2

    def copy_array(self, src: int, dst: int) -> None:
        check(load_library().e2d_copy_array(self._h, src, dst))

    def conserved_sums(self, which: int) -> np.ndarray:
        out = (C.c_double * 4)()
        check(load_library().e2d_conserved_sums(self._h, which, out))
        return np.array(out[:])

    def kinetic_cells(self, which: int) -> int:
        """Interior cells whose internal energy is below kin_frac of the total
        (in split precision: the cells converted again in FP64)."""
        out = C.c_longlong(0)
        check(load_library().e2d_kinetic_cells(self._h, which, C.byref(out)))
        return out.value

    def timers(self) -> dict:
        out = (C.c_double * 5)()
        check(load_library().e2d_timers(self._h, out))
        return dict(zip(("boundaries", "godunov", "primitives", "fluxes", "update"), out[:]))

    def save_vtk(self, which: int, output_dir: str | Path = ".", output_prefix: str = "output",
                 step: int | None = None) -> Path:
        """HydroRun::saveVTK: writes <output_dir>/<output_prefix>_<step:07d>.vti."""
        step = self.nStep if step is None else step
        check(load_library().e2d_save_vtk(self._h, which, str(output_dir).encode(),
                                          output_prefix.encode(), step))
        return Path(output_dir) / f"{output_prefix}_{step:07d}.vti"

    def radial_profile(self, nbins: int | None = None, which: int = _capi.E2D_U):
        """ComputeRadialProfileFunctor: (bin centres, mean density, cell counts)."""
        n = self.params.blast_nbins if nbins is None else nbins
        dist = np.empty(n)
        dens = np.empty(n)
        cnt = np.empty(n, dtype=np.int64)
        check(load_library().e2d_radial_profile(self._h, which, n, dist.ctypes.data,
                                                dens.ctypes.data, cnt.ctypes.data))
        return dist, dens, cnt


def write_vti(params, interior: np.ndarray, output_dir: str | Path = ".",
              output_prefix: str = "output", step: int = 0, float32: bool = False) -> Path:
    """The .vti file of HydroRun::saveVTK from a host array [var][j][i] (nx*ny)."""
    p = params.raw if isinstance(params, HydroParams) else params
    a = np.ascontiguousarray(interior, dtype=np.float64)
    if a.shape != (4, p.ny, p.nx):
        raise ValueError(f"expected shape (4, {p.ny}, {p.nx}), got {a.shape}")
    check(load_library().e2d_write_vti(C.byref(p), a.ctypes.data, int(float32),
                                       str(output_dir).encode(), output_prefix.encode(), step))
    return Path(output_dir) / f"{output_prefix}_{step:07d}.vti"


def write_npy(path: str | Path, data: np.ndarray) -> None:
    """cnpy::npy_save of a 1-D double array."""
    a = np.ascontiguousarray(data, dtype=np.float64).ravel()
    check(load_library().e2d_write_npy(str(path).encode(), a.ctypes.data, a.size))


def device_info() -> tuple[str, int]:
    buf = C.create_string_buffer(256)
    ncu = C.c_int(0)
    check(load_library().e2d_device_info(buf, 256, C.byref(ncu)))
    return buf.value.decode(), ncu.value
