"""Batched driver: many independent instances advanced together, in one
persistent launch per run (the default) or one launch per step.

HydroBatch wraps the e2d_batch_* C ABI.  sweep_params() builds the seeded
512-instance sweep: sizes nx, ny drawn independently from
{64, 96, 128, 192, 256, 384, 512}; even instances are implodes with the
high-state pressure uniform in [0.5, 2], odd ones blasts with disk radius
uniform in [0.05, 0.2] and pressure ratio log-uniform in [10, 1e4]
(p_out = 0.1, rho = 1); gamma 1.4, reflecting walls, HLLC, CFL 0.8.

    python -m euler2d_kokkos_amd.batch [--count 512] [--steps 1000]
        [--precision split|fp64|fp32] [--launch auto|persistent|steps] [--tile T]

runs the seeded sweep on the GPU and prints one JSON line (cell-updates per
second over the device time and over the host clock).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import E2DParams, check, load_library

SIZES = (64, 96, 128, 192, 256, 384, 512)
SWEEP_SEED = 20261202


def default_params() -> E2DParams:
    """HydroParams::setup on an empty deck (the reference's defaults)."""
    p = E2DParams()
    err = C.c_int(0)
    rc = load_library().e2d_params_from_ini(b"/nonexistent/empty.ini", C.byref(p), C.byref(err))
    if rc not in (_capi.E2D_OK, _capi.E2D_ERR_IO):
        check(rc)
    return p


def make_params(**fields) -> E2DParams:
    p = default_params()
    for k, v in fields.items():
        setattr(p, k, v)
    check(load_library().e2d_params_init(C.byref(p)))
    return p


def sweep_params(count: int = 512, steps: int = 1000, seed: int = SWEEP_SEED,
                 sizes=SIZES) -> list[E2DParams]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        nx = int(rng.choice(sizes))
        ny = int(rng.choice(sizes))
        common = dict(nx=nx, ny=ny, gamma0=1.4, cfl=0.8, nStepmax=steps, tEnd=1e9,
                      slope_type=1.0, riemannSolverType=2, honor_riemann=0,
                      boundary_type_xmin=1, boundary_type_xmax=1,
                      boundary_type_ymin=1, boundary_type_ymax=1)
        if k % 2 == 0:
            out.append(make_params(problemType=0, implode_pressure_in=float(rng.uniform(0.5, 2.0)),
                                   **common))
        else:
            r = float(rng.uniform(0.05, 0.2))
            ratio = 10.0 ** float(rng.uniform(1.0, 4.0))
            out.append(make_params(problemType=1, blast_radius=r, blast_center_x=0.5,
                                   blast_center_y=0.5, blast_density_in=1.0,
                                   blast_density_out=1.0, blast_pressure_out=0.1,
                                   blast_pressure_in=0.1 * ratio, **common))
    return out


class HydroBatch:
    def __init__(self, params: list[E2DParams], precision: int = _capi.E2D_PRECISION_FP32_SPLIT,
                 initialise: bool = True):
        arr = (E2DParams * len(params))(*params)
        self.params = list(params)
        self._h = C.c_void_p()
        check(load_library().e2d_batch_create(arr, len(params), precision, C.byref(self._h)))
        if initialise:
            self.init()

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().e2d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init(self) -> None:
        check(load_library().e2d_batch_init(self._h))

    def set_tile(self, seg_rows: int) -> None:
        check(load_library().e2d_batch_set_tile(self._h, seg_rows))

    def set_launch(self, mode: int, spin_limit: int = 0) -> None:
        """E2D_BATCH_LAUNCH_PERSISTENT (all the steps of a run in one launch;
        the default when the instances fill the XCD queues evenly) or
        E2D_BATCH_LAUNCH_STEPS (one launch per step, replayed from hipGraphs;
        the default otherwise); spin_limit bounds each dependency wait of the
        persistent launch."""
        check(load_library().e2d_batch_set_launch(self._h, mode, spin_limit))

    def get_launch(self) -> int:
        out = C.c_int(-1)
        check(load_library().e2d_batch_get_launch(self._h, C.byref(out)))
        return out.value

    def queue_info(self) -> tuple[int, float]:
        """(XCD queues of the persistent launch, fullest queue's cells over the mean)."""
        n = C.c_int(0)
        imb = C.c_double(0.0)
        check(load_library().e2d_batch_queue_info(self._h, C.byref(n), C.byref(imb)))
        return n.value, imb.value

    def run(self, steps: int, use_graph: bool = True) -> tuple[float, int]:
        """Runs `steps` batched steps; returns (device ms, instances still active)."""
        ms = C.c_double(0.0)
        act = C.c_int(0)
        check(load_library().e2d_batch_run(self._h, steps, int(use_graph), C.byref(ms),
                                           C.byref(act)))
        return ms.value, act.value

    def status(self, i: int) -> tuple[float, int]:
        t = C.c_double(0.0)
        n = C.c_int(0)
        check(load_library().e2d_batch_status(self._h, i, C.byref(t), C.byref(n)))
        return t.value, n.value

    def state(self, i: int, layout: int = _capi.E2D_LAYOUT_RIGHT) -> np.ndarray:
        p = self.params[i]
        shape = {_capi.E2D_LAYOUT_RIGHT: (p.isize, p.jsize, 4),
                 _capi.E2D_LAYOUT_LEFT: (4, p.jsize, p.isize),
                 _capi.E2D_LAYOUT_INTERIOR: (4, p.ny, p.nx)}[layout]
        out = np.empty(shape, dtype=np.float64)
        check(load_library().e2d_batch_copy_state(self._h, i, out.ctypes.data, layout))
        return out

    def info(self) -> dict:
        cells = C.c_longlong(0)
        waves = C.c_int(0)
        count = C.c_int(0)
        check(load_library().e2d_batch_info(self._h, C.byref(cells), C.byref(waves),
                                            C.byref(count)))
        return {"cells": cells.value, "waves": waves.value, "count": count.value}

    def autotune(self, candidates=(8, 10, 12, 14, 16, 20, 24, 32), steps: int = 5):
        """Times the candidate tiles (and, until set_launch fixes it, both launch
        modes where the queues allow the persistent one); keeps the fastest and
        returns (tile, ms per step)."""
        cand = (C.c_int * len(candidates))(*candidates)
        best = C.c_int(0)
        ms = C.c_double(0.0)
        check(load_library().e2d_batch_autotune(self._h, cand, len(candidates), steps,
                                                C.byref(best), C.byref(ms)))
        return best.value, ms.value


def density_l1(a: HydroBatch, b: HydroBatch) -> np.ndarray:
    """Relative L1 density error over the interior of every instance of batch
    `a` against the same instance of batch `b` (the same parameter list, e.g.
    a split-precision batch against its FP64 twin): sum|rho_a - rho_b| /
    sum|rho_b| per instance."""
    if len(a.params) != len(b.params):
        raise ValueError("the two batches hold different instance counts")
    out = np.empty(len(a.params))
    for i in range(len(a.params)):
        ra = a.state(i, _capi.E2D_LAYOUT_INTERIOR)[0]
        rb = b.state(i, _capi.E2D_LAYOUT_INTERIOR)[0]
        out[i] = np.abs(ra - rb).sum() / np.abs(rb).sum()
    return out


def main(argv=None) -> int:
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(prog="python -m euler2d_kokkos_amd.batch",
                                 description="Runs the seeded instance sweep on the GPU.")
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--precision", choices=("split", "fp64", "fp32"), default="split")
    ap.add_argument("--launch", choices=("auto", "persistent", "steps"), default="auto",
                    help="auto: the batch's own default")
    ap.add_argument("--tile", type=int, default=0, help="rows per wavefront (0: autotune)")
    args = ap.parse_args(argv)
    prec = {"split": _capi.E2D_PRECISION_FP32_SPLIT, "fp64": _capi.E2D_PRECISION_FP64,
            "fp32": _capi.E2D_PRECISION_FP32}[args.precision]
    ps = sweep_params(count=args.count, steps=args.steps)
    b = HydroBatch(ps, precision=prec)
    if args.launch != "auto":
        b.set_launch(_capi.E2D_BATCH_LAUNCH_PERSISTENT if args.launch == "persistent"
                     else _capi.E2D_BATCH_LAUNCH_STEPS)
    tile = args.tile
    if tile > 0:
        b.set_tile(tile)
    else:
        b.run(60)
        tile, _ = b.autotune(candidates=(12, 16, 24, 32, 48, 64, 96), steps=60)
        b.init()
    t0 = time.perf_counter()
    ms, active = b.run(args.steps)
    wall = time.perf_counter() - t0
    updates = sum(p.nx * p.ny * b.status(i)[1] for i, p in enumerate(ps))
    print(json.dumps({"instances": len(ps), "steps": args.steps, "precision": args.precision,
                      "launch": ("persistent" if b.get_launch() == _capi.E2D_BATCH_LAUNCH_PERSISTENT
                                 else "steps"), "tile": tile,
                      "device_ms": round(ms, 3),
                      "gcell_per_s_device": round(updates / (ms * 1e-3) / 1e9, 3),
                      "gcell_per_s_wall": round(updates / wall / 1e9, 3),
                      "still_active": active}))
    b.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
