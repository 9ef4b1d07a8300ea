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
one persistent launch (and, for comparison, by one launch per step), and
reports every instance's L1 density error
against the same sweep run in FP64; `configs` times the other decks (implode
256x256, four-quadrant 8192x8192, shocked bubble 16384x8192) over a bounded
number of steps in both precisions, with the split run's L1 error and drift.
`cpu_baseline` times the CPU oracle (a restatement of the reference's
implementation 0 with OpenMP) on this host: the median of three bounded
samples, with the host's CPU model, socket count and thread placement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

# OpenMP thread placement of the CPU baseline; read when the OpenMP runtime
# starts, so set before any library that links it is loaded
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

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


# VALU issue peak: one wave64 VALU instruction per SIMD every 4 cycles (16
# lanes per cycle), 1024 SIMDs at the 2.4 GHz peak engine clock; for FP64 FMAs
# this is the 78.6 TFLOP/s vector peak
VALU_PEAK_INSTS = 1024 * 2.4e9 / 4


def profiled_valu(workload: str, tile: int):
    """SQ_INSTS_VALU per launch of the FP64 step kernel from the committed
    rocprofv3 counter summary of this workload and tile, if there is one."""
    import re
    if workload != "blast":
        return None
    for path in sorted(PROFILES.glob("step_fused_fp64_blast_1024x1536*.md")):
        text = path.read_text()
        m = re.search(r"^valu_insts_per_launch: (\d+)", text, re.M)
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
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other decks (implode, four-quadrant, shocked bubble)")
    ap.add_argument("--config-steps", type=int, default=100,
                    help="steps of each other deck in the `configs` record")
    ap.add_argument("--workload", default="blast",
                    choices=["blast", "four_quadrant", "shocked_bubble", "implode"],
                    help="deck to time (blast 1024x1536 is the headline workload)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="K",
                    help="only launch the plain whole-array copy kernel K times (the known-size "
                         "reference for rocprof FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def batch_line(steps: int) -> dict:
    """The 512-instance ragged sweep (FP32 split, HLLC, 1000 steps each) in one
    persistent launch (per-XCD work queues); tile autotuned first.  The same
    sweep then runs with one launch per step, replayed from hipGraphs (tile
    autotuned for that mode), and both again in FP64; every instance's relative
    L1 density error against its FP64 twin is reported."""
    import numpy as np

    from euler2d_kokkos_amd import (E2D_BATCH_LAUNCH_PERSISTENT, E2D_BATCH_LAUNCH_STEPS,
                                    E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64)
    from euler2d_kokkos_amd.batch import HydroBatch, density_l1, sweep_params

    ps = sweep_params(steps=steps)
    cand = (12, 16, 24, 32, 48, 64, 96)

    def timed(b, mode):
        b.set_launch(mode)
        b.run(60)  # the engine clock settles before the candidates are timed
        # candidates differ by a few percent, which runs of 20 steps do not
        # resolve; 60 + 2 x 7 x 62 steps stay within the sweep's 1000
        tile, _ = b.autotune(candidates=cand, steps=60)
        b.init()
        t0 = time.perf_counter()
        ms, active = b.run(steps, use_graph=True)
        wall = time.perf_counter() - t0
        cell_updates = sum(p.nx * p.ny * b.status(i)[1] for i, p in enumerate(ps))
        return {"value": round(cell_updates / wall / 1e6, 2), "tile": tile,
                "device_ms": round(ms, 3),
                "instance_steps_done": sum(b.status(i)[1] for i in range(len(ps))),
                "still_active": active}

    out = {}
    runs = {}
    for name, prec in (("fp32_split", E2D_PRECISION_FP32_SPLIT), ("fp64", E2D_PRECISION_FP64)):
        b = HydroBatch(ps, precision=prec)
        info = b.info()
        per_step = timed(b, E2D_BATCH_LAUNCH_STEPS)
        out[name] = timed(b, E2D_BATCH_LAUNCH_PERSISTENT)  # the default, run last: its
        out[name]["per_step_launches"] = per_step         # state is the one compared
        runs[name] = b
    err = density_l1(runs["fp32_split"], runs["fp64"])
    worst = int(np.argmax(err))
    pw = ps[worst]
    for b in runs.values():
        b.close()
    split = out["fp32_split"]
    return {"value": split["value"], "unit": "Mcell-updates/s",
            "instances": info["count"], "cells": info["cells"], "steps_each": steps,
            "instance_steps_done": split["instance_steps_done"],
            "still_active": split["still_active"], "tile": split["tile"],
            "device_ms": split["device_ms"], "precision": "fp32_split",
            "launch": "persistent", "per_step_launches": split["per_step_launches"],
            "fp64": out["fp64"],
            "speedup_vs_fp64": round(split["value"] / out["fp64"]["value"], 3),
            "rel_l1_density_vs_fp64": {
                "max": float(err.max()), "median": float(np.median(err)),
                "mean": float(err.mean()), "instances_above_1e-4": int((err > 1e-4).sum()),
                "worst": {"instance": worst, "nx": pw.nx, "ny": pw.ny,
                          "problem": "implode" if pw.problemType == 0 else "blast",
                          "pressure_ratio": round(pw.blast_pressure_in / pw.blast_pressure_out
                                                  if pw.problemType == 1 else
                                                  pw.implode_pressure_in / 0.14, 2),
                          "rel_l1": float(err[worst])}}}


def host_cpu() -> dict:
    """CPU model, sockets and logical CPUs of this host (/proc/cpuinfo)."""
    model, sockets = "unknown", set()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "sockets": max(len(sockets), 1), "nproc": os.cpu_count() or 1,
            "usable_cpus": usable}


def cpu_baseline(deck: Path, seconds: float) -> dict:
    """Times the CPU oracle on this host: the median of three samples of the
    workload, each bounded to a third of `seconds`.  Threads: OMP_NUM_THREADS
    when set (the CPU share a GPU job gets on a shared host), else every CPU
    this process may run on; placement OMP_PROC_BIND / OMP_PLACES."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # test infrastructure, used here only as the baseline

    host = host_cpu()
    threads = int(os.environ.get("E2D_CPU_THREADS") or os.environ.get("OMP_NUM_THREADS")
                  or host["usable_cpus"])
    O.lib().ora_set_threads(threads)
    p = O.params_from_ini(deck)
    run = O.OracleRun(p)
    run.run(1)  # warm-up step (first touch of the arrays)
    rates, counts, spans = [], [], []
    for _ in range(3):
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds / 3 and run.nStep < p.nStepmax:
            run.run(1)
            steps += 1
        el = time.perf_counter() - t0
        rates.append(p.nx * p.ny * steps / el / 1e6)
        counts.append(steps)
        spans.append(round(el, 2))
    value = sorted(rates)[1]
    return {"value": round(value, 3), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "samples": [round(r, 3) for r in rates],
            "host": host, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES"),
            "sample": f"{deck.name}: median of 3 samples of {counts} steps ({spans} s) after 1 "
                      f"warm-up step, CPU oracle (FP64) with OpenMP over {threads} threads"}


def rel_l1_runs(run, ref) -> list:
    """Per-variable relative L1 distance of two runs' current interiors, one
    variable at a time (sum|a - b| / sum|b|; the absolute sum where the
    reference variable is zero everywhere, as a y velocity at t = 0)."""
    import numpy as np

    from euler2d_kokkos_amd import E2D_LAYOUT_INTERIOR

    a = run.get_state(run.current(), E2D_LAYOUT_INTERIOR)
    b = ref.get_state(ref.current(), E2D_LAYOUT_INTERIOR)
    out = []
    for v in range(4):
        num = float(np.abs(a[v] - b[v]).sum())
        den = float(np.abs(b[v]).sum())
        out.append(num / den if den > 0 else num)
    return out


# GPU time of the untimed settle run before a timed region: the engine clock
# climbs over the first ~10 ms of sustained work (20-step windows timed right
# after the tile search ran 4% slower than later ones)
SETTLE_MS = 20.0


def settle(run, ms_per_step):
    """Keeps the GPU busy on the chosen tile for about SETTLE_MS before the
    warm-up (the state is re-initialised before the timed region)."""
    n = int(min(400, max(30, SETTLE_MS / max(ms_per_step, 1e-3))))
    run.run(n)
    run.synchronize()
    return n


def timed_run(params, precision, steps, tile=0):
    """A fresh run of the deck: tile autotuned on the initial state (or fixed),
    a settle run and 30 warm-up steps (the run loop's graphs are captured
    there), then `steps` steps from t = 0 timed with the host clock and with
    HIP events."""
    from euler2d_kokkos_amd import E2D_U, HydroRun

    run = HydroRun(params, precision=precision)
    if tile > 0:
        run.set_tile(tile)
        ms = run.time_loop(5)[0] / 5  # the settle run's length comes from it
    else:
        _, ms = run.autotune(steps=10)
    settle(run, ms)
    run.init_problem()
    run.run(30)
    run.synchronize()
    run.init_problem()
    s0 = run.conserved_sums(E2D_U)
    t0 = time.perf_counter()
    ms, executed = run.time_loop(steps)
    run.synchronize()
    wall = time.perf_counter() - t0
    return run, s0, ms, executed, wall


def split_comparison(params, run64, steps, cells, fp64_ms, tile=0) -> dict:
    """Same workload in FP32 split precision: throughput, L1 error against the
    FP64 run's final state, FP64-summed drift of mass and energy over the run
    (nonzero where the boundaries let mass through), and the distance of the
    final mass and energy from the FP64 run's."""
    from euler2d_kokkos_amd import E2D_PRECISION_FP32_SPLIT

    run, s0, ms, executed, wall = timed_run(params, E2D_PRECISION_FP32_SPLIT, steps, tile)
    l1 = rel_l1_runs(run, run64)
    s1 = run.conserved_sums(run.current())
    s64 = run64.conserved_sums(run64.current())
    value = cells * executed / wall / 1e6
    per_step = ms / executed
    out = {"value": round(value, 2), "unit": "Mcell-updates/s",
           "speedup_vs_fp64": round(fp64_ms / per_step, 3),
           "kernel_ms_per_step": round(per_step, 5),
           "roofline_frac": round(BYTES_PER_CELL * cells / (per_step * 1e-3) / 1e9
                                  / HBM_PEAK_GBS, 4),
           "rel_l1_vs_fp64": l1,
           "mass_drift": float(abs(s1[0] - s0[0]) / abs(s0[0])),
           "energy_drift": float(abs(s1[1] - s0[1]) / abs(s0[1])),
           "mass_vs_fp64": float(abs(s1[0] - s64[0]) / abs(s64[0])),
           "energy_vs_fp64": float(abs(s1[1] - s64[1]) / abs(s64[1])),
           "kinetic_cells": run.kinetic_cells(run.current()),
           "tile": run.get_tile(), "final_t": run.t}
    run.close()
    return out


def config_pair(workload: str, steps: int) -> dict:
    """One of the other decks in FP64 and in split precision over `steps`
    steps from t = 0 (the deck's own count is larger; the grid is full size):
    throughput of each, the split run's L1 error against the FP64 run, and its
    mass / energy drift."""
    from euler2d_kokkos_amd import E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64, HydroParams

    deck = ROOT / "euler2d_kokkos_amd" / "decks" / DECKS[workload][0]
    params = HydroParams.from_ini(deck)
    steps = min(steps, params.nStepmax)
    cells = params.nx * params.ny
    r64, _, ms64, ex64, wall64 = timed_run(params, E2D_PRECISION_FP64, steps)
    r32, s0, ms32, ex32, wall32 = timed_run(params, E2D_PRECISION_FP32_SPLIT, steps)
    if ex64 != steps or ex32 != steps:
        raise RuntimeError(f"{workload}: ran {ex64} / {ex32} steps instead of {steps}")
    s1 = r32.conserved_sums(r32.current())
    l1 = rel_l1_runs(r32, r64)
    k64, k32 = ms64 / steps, ms32 / steps
    out = {"deck": DECKS[workload][0], "nx": params.nx, "ny": params.ny, "steps": steps,
           "deck_steps": DECKS[workload][2],
           "fp64": {"value": round(cells * steps / wall64 / 1e6, 2),
                    "kernel_ms_per_step": round(k64, 5), "tile": r64.get_tile(),
                    "roofline_frac": round(BYTES_PER_CELL * cells / (k64 * 1e-3) / 1e9
                                           / HBM_PEAK_GBS, 4)},
           "fp32_split": {"value": round(cells * steps / wall32 / 1e6, 2),
                          "kernel_ms_per_step": round(k32, 5), "tile": r32.get_tile(),
                          "roofline_frac": round(BYTES_PER_CELL * cells / (k32 * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 4)},
           "speedup_split_vs_fp64": round(k64 / k32, 3),
           "rel_l1_split_vs_fp64": l1,
           "split_mass_drift": float(abs(s1[0] - s0[0]) / abs(s0[0])),
           "split_energy_drift": float(abs(s1[1] - s0[1]) / abs(s0[1])),
           "split_kinetic_cells": r32.kinetic_cells(r32.current())}
    r32.close()
    r64.close()
    return out


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
    tune_ms = 1.0
    if args.tile > 0:
        run.set_tile(args.tile)
        tune_ms = run.time_loop(5)[0] / 5  # the settle run's length comes from it
    elif args.calibrate_copy <= 0:
        _, tune_ms = run.autotune(steps=20)

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

    settle_steps = settle(run, tune_ms)
    run.init_problem()
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
    # the second roof: VALU instructions per launch (committed counter summary
    # of this kernel, workload and tile) issued in this run's kernel time,
    # against one wave64 instruction per SIMD every 4 cycles
    valu_insts = profiled_valu(args.workload, run.get_tile())
    valu_frac = (round(valu_insts / (per_step_ms * 1e-3) / VALU_PEAK_INSTS, 4)
                 if valu_insts else None)
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
                   "tile": run.get_tile(), "device": dev, "cus": ncu,
                   "untimed_setup": f"tile search, then {settle_steps} settle steps on the "
                                    f"chosen tile (engine clock), then the {args.warmup} "
                                    "warm-up steps; the state is re-initialised after each"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(args.workload, run.get_tile()),
                     "valu_frac": valu_frac, "valu_insts_per_launch": valu_insts},
        "kernel_ms_per_step": round(per_step_ms, 5),
        "final_t": run.t,
    }
    if not args.no_split:
        line["fp32_split"] = split_comparison(params, run, args.steps, cells, per_step_ms,
                                              args.split_tile)
    if not args.no_batch:
        line["batched"] = batch_line(args.batch_steps)
    if not args.no_configs:
        line["configs"] = {w: config_pair(w, args.config_steps)
                           for w in ("implode", "four_quadrant", "shocked_bubble")
                           if w != args.workload}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(deck, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
