"""Benchmark of the unsplit MUSCL-Hancock Godunov timestep on one MI355X.

Workload: the blast problem on 1024x1536 cells ([0,1]x[0,1.5], gamma 1.4,
p=10 in a disk of radius 0.1 at (0.5,0.75), p=0.1 outside, reflecting walls,
HLLC, CFL 0.7), FP64, 2000 steps from t=0 -- one step is one timestep of the
whole grid (dt reduction included, dt kept on the device).  The state lives in
HBM before the timed region starts; nothing crosses PCIe inside it.

  python bench.py [--gpus 1] [--steps 2000] [--warmup 50] [--workload blast]

The step kernel's tile (rows marched per wavefront) is autotuned on the initial
state before the timed region unless --tile fixes it.  Prints one JSON line.  `value` is interior cell-updates per second over the
timed region (host clock, stream synchronised on both sides).  `roofline` uses
the step kernel's average duration from HIP events on the solver's stream and
the algorithmic 64 bytes per cell-update (read 4 doubles, write 4 doubles).
`fp32_split` repeats the run in split precision (hi+lo FP32 state) and reports
its distance from the FP64 result; `batched` times the seeded 512-instance
ragged sweep (sizes 64..512 per side, FP32 split, 1000 steps each) advanced by
one batched launch per step.  `cpu_baseline` times the CPU oracle (a restatement of the reference's
implementation 0 with OpenMP) on this host over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s on 1 MI355X: FP64 plain vs FP32/batched; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 64.0
DECKS = {  # deck, description, steps of the workload
    "blast": ("blast_1024x1536.ini", "blast 1024x1536 FP64 HLLC CFL 0.7 reflecting walls", 2000),
    "four_quadrant": ("four_quadrant_8192.ini", "four-quadrant 8192x8192 FP64 HLL CFL 0.8 "
                                                "transmissive", 1000),
    "shocked_bubble": ("shocked_bubble_16384x8192.ini", "shocked bubble 16384x8192 FP64 HLLC "
                                                        "CFL 0.8", 3000),
    "implode": ("implode_256.ini", "implode 256x256 FP64 HLLC CFL 0.8 reflecting walls", 500),
}


PROFILES = ROOT / "profiles"


def profiled_traffic(workload: str, tile: int):
    """HBM bytes per launch of the FP64 step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (one per tile), when one covers this
    workload and tile."""
    import re
    if workload != "blast":
        return None
    for path in sorted(PROFILES.glob("step_fused_fp64_blast_1024x1536*.md")):
        text = path.read_text()
        m = re.search(r"^traffic_bytes_per_launch: (\d+)", text, re.M)
        t = re.search(r"seg_rows (\d+)", text)
        if m and t and int(t.group(1)) == tile:
            return int(m.group(1))
    return None


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default: the workload's own count)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--split-tile", type=int, default=0,
                    help="tile of the split-precision run (default: autotuned)")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the FP32 split-precision comparison run")
    ap.add_argument("--tile", type=int, default=0,
                    help="rows marched by one wavefront of the step kernel (0: autotune)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 512-instance sweep")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--workload", default="blast",
                    choices=["blast", "four_quadrant", "shocked_bubble", "implode"],
                    help="deck to time (blast 1024x1536 is the headline workload)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="K",
                    help="only launch the plain whole-array copy kernel K times (the known-size "
                         "reference for rocprof FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def batch_line(steps: int) -> dict:
    """The 512-instance ragged sweep (FP32 split, HLLC, 1000 steps each) in one
    batched launch per step, replayed from hipGraphs; tile autotuned first."""
    from euler2d_kokkos_amd import E2D_PRECISION_FP32_SPLIT
    from euler2d_kokkos_amd.batch import HydroBatch, sweep_params

    ps = sweep_params(steps=steps)
    b = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    tile, _ = b.autotune(candidates=(8, 12, 16, 24, 32, 48, 64), steps=5)
    b.init()
    info = b.info()
    t0 = time.perf_counter()
    ms, active = b.run(steps, use_graph=True)
    wall = time.perf_counter() - t0
    done = sum(b.status(i)[1] for i in range(info["count"]))
    cell_updates = sum(p.nx * p.ny * b.status(i)[1] for i, p in enumerate(ps))
    b.close()
    return {"value": round(cell_updates / wall / 1e6, 2), "unit": "Mcell-updates/s",
            "instances": info["count"], "cells": info["cells"], "steps_each": steps,
            "instance_steps_done": done, "still_active": active, "tile": tile,
            "device_ms": round(ms, 3), "precision": "fp32_split", "launch": "hipGraph"}


def cpu_baseline(deck: Path, seconds: float) -> dict:
    """Times the CPU oracle on this host over a bounded sample of the workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # test infrastructure, used here only as the baseline

    threads = int(os.environ.get("E2D_CPU_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    O.lib().ora_set_threads(threads)
    p = O.params_from_ini(deck)
    run = O.OracleRun(p)
    run.run(1)  # warm-up step (first touch of the arrays)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and run.nStep < p.nStepmax:
        run.run(1)
        steps += 1
    el = time.perf_counter() - t0
    value = p.nx * p.ny * steps / el / 1e6
    return {"value": round(value, 3), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{deck.name}: {steps} steps after 1 warm-up step, CPU oracle "
                      f"(FP64) with OpenMP over {threads} threads, {el:.1f} s"}


def split_comparison(params, run64, steps, cells, fp64_ms, tile=0) -> dict:
    """Same workload in FP32 split precision: throughput, L1 error against the
    FP64 run's final state, FP64-summed drift of mass and energy over the run
    (nonzero where the boundaries let mass through), and the distance of the
    final mass and energy from the FP64 run's."""
    import numpy as np

    from euler2d_kokkos_amd import E2D_PRECISION_FP32_SPLIT, E2D_U, HydroRun

    run = HydroRun(params, precision=E2D_PRECISION_FP32_SPLIT)
    if tile > 0:
        run.set_tile(tile)
    else:
        run.autotune(steps=20)
    run.run(20)
    run.synchronize()
    run.init_problem()
    s0 = run.conserved_sums(E2D_U)
    t0 = time.perf_counter()
    ms, executed = run.time_loop(steps)
    run.synchronize()
    wall = time.perf_counter() - t0
    a = run.get_state(run.current())[2:-2, 2:-2, :].reshape(-1, 4)
    b = run64.get_state(run64.current())[2:-2, 2:-2, :].reshape(-1, 4)
    num = np.abs(a - b).sum(0)
    den = np.abs(b).sum(0)
    # a variable that is zero everywhere in the FP64 result (no y velocity yet)
    # reports its absolute L1 difference
    l1 = np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
    s1 = run.conserved_sums(run.current())
    s64 = run64.conserved_sums(run64.current())
    value = cells * executed / wall / 1e6
    per_step = ms / executed
    return {"value": round(value, 2), "unit": "Mcell-updates/s",
            "speedup_vs_fp64": round(fp64_ms / per_step, 3),
            "kernel_ms_per_step": round(per_step, 5),
            "roofline_frac": round(BYTES_PER_CELL * cells / (per_step * 1e-3) / 1e9
                                   / HBM_PEAK_GBS, 4),
            "rel_l1_vs_fp64": [float(x) for x in l1],
            "mass_drift": float(abs(s1[0] - s0[0]) / abs(s0[0])),
            "energy_drift": float(abs(s1[1] - s0[1]) / abs(s0[1])),
            "mass_vs_fp64": float(abs(s1[0] - s64[0]) / abs(s64[0])),
            "energy_vs_fp64": float(abs(s1[1] - s64[1]) / abs(s64[1])),
            "tile": run.get_tile(), "final_t": run.t}


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU "
                          "domain decomposition is a design note, not a code path",
                          "n_gpus": args.gpus}))
        return 2

    from euler2d_kokkos_amd import (E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64,
                                    E2D_SCHEDULE_FUSED, HydroParams, HydroRun, device_info)

    dev, ncu = device_info()
    if args.steps <= 0:
        args.steps = DECKS[args.workload][2]
    deck = ROOT / "euler2d_kokkos_amd" / "decks" / DECKS[args.workload][0]
    params = HydroParams.from_ini(deck)
    params.nStepmax = max(params.nStepmax, args.steps)
    run = HydroRun(params, schedule=E2D_SCHEDULE_FUSED)
    if args.tile > 0:
        run.set_tile(args.tile)
    elif args.calibrate_copy <= 0:
        run.autotune(steps=20)

    if args.calibrate_copy > 0:
        from euler2d_kokkos_amd import E2D_U, E2D_U2
        pitch = (params.isize + 15) // 16 * 16
        plane = (pitch * params.jsize + 63) // 64 * 64
        for _ in range(args.calibrate_copy):
            run.copy_array(E2D_U, E2D_U2)
        run.synchronize()
        print(json.dumps({"copy_kernel_launches": args.calibrate_copy,
                          "bytes_read_per_launch": plane * 4 * 8,
                          "bytes_written_per_launch": plane * 4 * 8}), flush=True)
        return 0

    if args.warmup > 0:
        run.run(args.warmup)
        run.synchronize()
    run.init_problem()  # the timed run starts from t = 0 like the workload
    run.synchronize()

    t0 = time.perf_counter()
    kernel_ms, executed = run.time_loop(args.steps)
    run.synchronize()
    wall = time.perf_counter() - t0
    if executed != args.steps:
        raise RuntimeError(f"ran {executed} steps instead of {args.steps}")

    cells = params.nx * params.ny
    value = cells * args.steps / wall / 1e6
    per_step_ms = kernel_ms / args.steps
    achieved = BYTES_PER_CELL * cells / (per_step_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic ({args.workload} initial condition generated on the GPU)",
        "config": {"workload": f"{DECKS[args.workload][1]}, {args.steps} steps from t=0",
                   "nx": params.nx, "ny": params.ny, "precision": "fp64", "schedule": "fused",
                   "tile": run.get_tile(), "device": dev, "cus": ncu},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(args.workload, run.get_tile())},
        "kernel_ms_per_step": round(per_step_ms, 5),
        "final_t": run.t,
    }
    if not args.no_split:
        line["fp32_split"] = split_comparison(params, run, args.steps, cells, per_step_ms,
                                              args.split_tile)
    if not args.no_batch:
        line["batched"] = batch_line(args.batch_steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(deck, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
