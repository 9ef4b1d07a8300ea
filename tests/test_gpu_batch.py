"""Batched driver on the GPU: every instance of a batch must equal the same
problem run alone (bit for bit: the batched kernel runs the same march per
wavefront), in FP64 (and hence the CPU oracle), in split precision and in the
plain FP32 comparator."""
import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import (E2D_BATCH_LAUNCH_PERSISTENT, E2D_BATCH_LAUNCH_STEPS,
                                E2D_PRECISION_FP32, E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64,
                                HydroParams, HydroRun)
from euler2d_kokkos_amd.batch import HydroBatch, density_l1, make_params, sweep_params

pytestmark = pytest.mark.gpu


def mixed_instances():
    return [
        make_params(nx=64, ny=96, problemType=0, cfl=0.8, nStepmax=30, tEnd=1e9),
        make_params(nx=128, ny=64, problemType=1, cfl=0.7, nStepmax=25, tEnd=1e9,
                    blast_radius=0.15, blast_pressure_in=50.0),
        make_params(nx=96, ny=96, problemType=0, cfl=0.8, nStepmax=1000, tEnd=0.02,
                    implode_pressure_in=1.7),
        make_params(nx=61, ny=130, problemType=1, cfl=0.6, nStepmax=40, tEnd=1e9,
                    boundary_type_xmin=3, boundary_type_xmax=3, blast_center_y=0.3),
        make_params(nx=200, ny=33, problemType=3, cfl=0.8, nStepmax=20, tEnd=1e9,
                    boundary_type_ymin=2, boundary_type_ymax=2),
        make_params(nx=64, ny=64, problemType=2, cfl=0.8, nStepmax=35, tEnd=1e9,
                    boundary_type_xmin=2, boundary_type_xmax=2, boundary_type_ymin=2,
                    boundary_type_ymax=2),
    ]


@pytest.mark.parametrize("precision", [E2D_PRECISION_FP64, E2D_PRECISION_FP32_SPLIT,
                                       E2D_PRECISION_FP32], ids=["fp64", "split", "fp32"])
@pytest.mark.parametrize("launch", ["launches", "graph", "persistent"])
def test_batch_equals_single_runs(gpu_available, precision, launch):
    ps = mixed_instances()
    batch = HydroBatch(ps, precision=precision)
    batch.set_launch(E2D_BATCH_LAUNCH_PERSISTENT if launch == "persistent"
                     else E2D_BATCH_LAUNCH_STEPS)
    batch.set_tile(12)
    ms, active = batch.run(64, use_graph=launch == "graph")
    assert active == 0
    for i, p in enumerate(ps):
        run = HydroRun(HydroParams(p.copy()), precision=precision)
        run.set_tile(7)  # results do not depend on the tile
        run.run()
        t, n = batch.status(i)
        assert (t, n) == (run.t, run.nStep), i
        np.testing.assert_array_equal(batch.state(i), run.get_state(run.current()),
                                      err_msg=f"instance {i}")


def test_batch_graph_replays_from_any_phase(gpu_available):
    """Runs of uneven lengths, with and without the captured graph, leave the
    batch where single runs end.  The loop control cycles through three ring
    entries and the graph replays only from the first one, so each call below
    starts the graph from a different launch count (7, 38, then 103)."""
    ps = mixed_instances()
    for p in ps:  # every instance still active past the graph replays
        p.nStepmax = 150
        p.tEnd = 1e9
    batch = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    batch.set_launch(E2D_BATCH_LAUNCH_STEPS)
    batch.set_tile(12)
    steps = 0
    for k, g in ((7, False), (31, True), (65, True), (2, True)):
        batch.run(k, use_graph=g)
        steps += k
        for i in (0, 3):
            assert batch.status(i)[1] == steps, (i, steps)
    for i, p in enumerate(ps):
        q = p.copy()
        q.nStepmax = steps
        run = HydroRun(HydroParams(q), precision=E2D_PRECISION_FP32_SPLIT)
        run.run()
        assert batch.status(i) == (run.t, run.nStep), i
        np.testing.assert_array_equal(batch.state(i), run.get_state(run.current()),
                                      err_msg=f"instance {i}")


@pytest.mark.parametrize("precision", [E2D_PRECISION_FP64, E2D_PRECISION_FP32_SPLIT],
                         ids=["fp64", "split"])
@pytest.mark.parametrize("launch", [E2D_BATCH_LAUNCH_STEPS, E2D_BATCH_LAUNCH_PERSISTENT],
                         ids=["steps", "persistent"])
def test_batch_wide_addressing(gpu_available, precision, launch, monkeypatch):
    """A batch with the 64-bit row addressing (forced) equals single runs with
    the 32-bit one."""
    ps = mixed_instances()
    monkeypatch.setenv("E2D_WIDE_ADDRESSING", "1")
    batch = HydroBatch(ps, precision=precision)
    monkeypatch.delenv("E2D_WIDE_ADDRESSING")
    batch.set_launch(launch)
    batch.set_tile(16)
    batch.run(64)
    for i, p in enumerate(ps):
        run = HydroRun(HydroParams(p.copy()), precision=precision)
        run.run()
        assert batch.status(i) == (run.t, run.nStep), i
        np.testing.assert_array_equal(batch.state(i), run.get_state(run.current()),
                                      err_msg=f"instance {i}")


@pytest.mark.parametrize("slopes", ["minmod", "mixed"])
@pytest.mark.parametrize("launch", [E2D_BATCH_LAUNCH_STEPS, E2D_BATCH_LAUNCH_PERSISTENT],
                         ids=["steps", "persistent"])
def test_batch_fp64_equals_oracle(gpu_available, launch, slopes):
    """Every instance's state and t equal the oracle's bit for bit.  "minmod":
    all instances have slope_type 1, so the batch runs its minmod kernels;
    "mixed": one instance has the MC limiter (slope_type 2), so the batch runs
    the generic ones."""
    ps = mixed_instances()[:3]
    if slopes == "mixed":
        ps[1] = make_params(nx=128, ny=64, problemType=1, cfl=0.7, nStepmax=25, tEnd=1e9,
                            blast_radius=0.15, blast_pressure_in=50.0, slope_type=2.0)
    batch = HydroBatch(ps, precision=E2D_PRECISION_FP64)
    batch.set_launch(launch)
    batch.run(40)
    for i, p in enumerate(ps):
        ora = O.OracleRun(p)
        ora.run()
        assert batch.status(i) == (ora.t, ora.nStep)
        np.testing.assert_array_equal(batch.state(i), ora.current())


def test_batch_run_requires_init(gpu_available):
    """A batch whose instances have no initial state and no startup dt yet
    refuses to step, then runs once e2d_batch_init has set them."""
    from euler2d_kokkos_amd._capi import E2DError
    ps = mixed_instances()[:2]
    batch = HydroBatch(ps, precision=E2D_PRECISION_FP64, initialise=False)
    with pytest.raises(E2DError, match="e2d_batch_init"):
        batch.run(3)
    with pytest.raises(E2DError, match="e2d_batch_init"):
        batch.autotune(candidates=(8,), steps=1)
    batch.init()
    batch.run(3)
    assert batch.status(0)[1] == 3


def test_batch_autotune_and_reinit(gpu_available):
    ps = mixed_instances()
    batch = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    best, ms = batch.autotune(candidates=(6, 12, 24), steps=3)
    assert best in (6, 12, 24) and ms > 0
    batch.init()
    batch.run(10, use_graph=False)
    a = [batch.state(i) for i in range(len(ps))]
    batch.init()
    batch.set_tile(9)
    batch.run(10, use_graph=True)
    for i in range(len(ps)):
        np.testing.assert_array_equal(batch.state(i), a[i])


def test_sweep_instances_accuracy(gpu_available):
    """32 instances of the seeded 512-instance sweep, chosen to include its
    strongest blasts (pressure ratio >= 1e3) and its 512-cell sides, run 1000
    steps in split precision and in FP64: every instance's relative L1 density
    error against its FP64 twin stays within 1e-4."""
    ps = sweep_params(steps=1000)
    ratio = [p.blast_pressure_in / p.blast_pressure_out if p.problemType == 1 else 0.0
             for p in ps]
    strong = sorted(range(len(ps)), key=lambda i: -ratio[i])[:12]
    wide = [i for i in range(len(ps)) if 512 in (ps[i].nx, ps[i].ny) and i not in strong][:10]
    rest = [i for i in range(0, len(ps), 37) if i not in strong and i not in wide]
    pick = (strong + wide + rest)[:32]
    sel = [ps[i] for i in pick]
    assert min(ratio[i] for i in strong) >= 1e3
    assert any(p.nx == 512 and p.ny == 512 for p in sel) or len(wide) >= 5
    b64 = HydroBatch(sel, precision=E2D_PRECISION_FP64)
    b32 = HydroBatch(sel, precision=E2D_PRECISION_FP32_SPLIT)
    for b in (b64, b32):
        _, active = b.run(1000)
        assert active == 0
    for i in range(len(sel)):
        assert b32.status(i)[1] == b64.status(i)[1] == 1000
    err = density_l1(b32, b64)
    worst = int(np.argmax(err))
    print(f"sweep subset: instances={len(sel)} max={err.max():.3e} median={np.median(err):.3e} "
          f"worst=sweep[{pick[worst]}] {sel[worst].nx}x{sel[worst].ny} "
          f"problem={sel[worst].problemType} ratio={ratio[pick[worst]]:.0f}", flush=True)
    assert np.all(err <= 1e-4), err


@pytest.mark.parametrize("precision", [E2D_PRECISION_FP64, E2D_PRECISION_FP32_SPLIT,
                                       E2D_PRECISION_FP32], ids=["fp64", "split", "fp32"])
def test_persistent_launch_equals_single_runs(gpu_available, precision):
    """All the steps of a run in one persistent launch (per-XCD work queues,
    per-instance dependency counts) leave every instance where its single run
    ends, bit for bit: instances that finish early (nStepmax, tEnd) stay put,
    and two runs in a row continue the control rings' phases."""
    ps = mixed_instances()
    batch = HydroBatch(ps, precision=precision)
    batch.set_tile(12)
    batch.set_launch(E2D_BATCH_LAUNCH_PERSISTENT)
    assert batch.get_launch() == E2D_BATCH_LAUNCH_PERSISTENT
    batch.run(7)
    ms, active = batch.run(57)
    assert active == 0 and ms > 0
    for i, p in enumerate(ps):
        run = HydroRun(HydroParams(p.copy()), precision=precision)
        run.run()
        t, n = batch.status(i)
        assert (t, n) == (run.t, run.nStep), i
        np.testing.assert_array_equal(batch.state(i), run.get_state(run.current()),
                                      err_msg=f"instance {i}")


def test_persistent_launch_sweep_equals_step_launches(gpu_available):
    """64 instances of the seeded sweep, 120 steps: the persistent launch and
    the per-step launches give the same bits for every instance."""
    ps = sweep_params(count=64, steps=120)
    a = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    b = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    for x in (a, b):
        x.set_tile(16)
    a.set_launch(E2D_BATCH_LAUNCH_STEPS)
    b.set_launch(E2D_BATCH_LAUNCH_PERSISTENT)
    a.run(120)
    b.run(120)
    for i in range(len(ps)):
        assert a.status(i) == b.status(i), i
        np.testing.assert_array_equal(a.state(i), b.state(i), err_msg=f"instance {i}")


def test_persistent_launch_wait_limit(gpu_available):
    """A dependency wait bounded to one sleep round runs out on a large
    instance (its next step's items are taken while its step is still
    marching): the launch ends, e2d_batch_run reports it, and the batch runs
    normally again after e2d_batch_init."""
    from euler2d_kokkos_amd import E2DError
    ps = [make_params(nx=512, ny=512, problemType=1, cfl=0.7, nStepmax=100, tEnd=1e9)]
    batch = HydroBatch(ps, precision=E2D_PRECISION_FP64)
    batch.set_tile(8)
    batch.set_launch(E2D_BATCH_LAUNCH_PERSISTENT, spin_limit=1)
    with pytest.raises(E2DError):
        batch.run(20)
    batch.set_launch(E2D_BATCH_LAUNCH_STEPS)
    batch.init()
    _, active = batch.run(20)
    assert active == 1 and batch.status(0)[1] == 20


def test_default_launch_follows_queue_balance(gpu_available):
    """The seeded 512-instance sweep fills the XCD queues evenly and takes the
    persistent launch; six grids of different sizes, or one grid, would leave
    XCDs idle and take one launch per step."""
    b = HydroBatch(sweep_params(steps=10), precision=E2D_PRECISION_FP32_SPLIT)
    nxcd, imbalance = b.queue_info()
    print(f"XCD queues {nxcd}, sweep imbalance {imbalance:.4f}", flush=True)
    assert nxcd >= 1 and 1.0 <= imbalance <= 1.10
    assert b.get_launch() == E2D_BATCH_LAUNCH_PERSISTENT
    b.close()
    for ps in (mixed_instances(), mixed_instances()[:1]):
        b = HydroBatch(ps, precision=E2D_PRECISION_FP64)
        n, imb = b.queue_info()
        expect = E2D_BATCH_LAUNCH_PERSISTENT if imb <= 1.10 else E2D_BATCH_LAUNCH_STEPS
        assert b.get_launch() == expect, (n, imb)
        if n > 1 and len(ps) == 1:  # one grid on one of several XCDs
            assert b.get_launch() == E2D_BATCH_LAUNCH_STEPS and imb == n
        b.close()


def test_autotune_picks_launch_mode_and_tile(gpu_available):
    """With the launch mode left to the batch and even queues, the autotuner
    times both modes and keeps the fastest (mode, tile) pair; a mode fixed by
    set_launch stays.  Either way the batch then runs to the same bits."""
    ps = sweep_params(count=64, steps=40)
    a = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    assert a.queue_info()[1] <= 1.10
    tile, ms = a.autotune(candidates=(16, 24), steps=3)
    assert tile in (16, 24) and ms > 0
    assert a.get_launch() in (E2D_BATCH_LAUNCH_PERSISTENT, E2D_BATCH_LAUNCH_STEPS)
    b = HydroBatch(ps, precision=E2D_PRECISION_FP32_SPLIT)
    b.set_launch(E2D_BATCH_LAUNCH_STEPS)
    b.autotune(candidates=(16, 24), steps=3)
    assert b.get_launch() == E2D_BATCH_LAUNCH_STEPS
    for x in (a, b):
        x.init()
        x.run(40)
    for i in range(len(ps)):
        assert a.status(i) == b.status(i), i
        np.testing.assert_array_equal(a.state(i), b.state(i), err_msg=f"instance {i}")


def test_persistent_run_split_over_launches(gpu_available, monkeypatch):
    """A run whose item count passes the per-launch limit (lowered here with
    E2D_PERSISTENT_MAX_ITEMS to 500 items)
    takes several persistent launches; the control rings' phases continue
    across them, and the instances end where per-step launches leave them."""
    ps = sweep_params(count=16, steps=50)
    a = HydroBatch(ps, precision=E2D_PRECISION_FP64)
    a.set_launch(E2D_BATCH_LAUNCH_STEPS)
    monkeypatch.setenv("E2D_PERSISTENT_MAX_ITEMS", "500")
    b = HydroBatch(ps, precision=E2D_PRECISION_FP64)
    monkeypatch.delenv("E2D_PERSISTENT_MAX_ITEMS")
    b.set_launch(E2D_BATCH_LAUNCH_PERSISTENT)
    for x in (a, b):
        x.set_tile(12)
    # the fullest XCD queue holds at least waves / nxcd items per step, so a
    # launch covers at most 500 * nxcd / waves steps: fewer than 11
    nxcd, _ = b.queue_info()
    assert b.info()["waves"] * 11 > 500 * nxcd
    a.run(11)
    a.run(39)
    b.run(11)
    b.run(39)
    for i in range(len(ps)):
        assert a.status(i) == b.status(i), i
        np.testing.assert_array_equal(a.state(i), b.state(i), err_msg=f"instance {i}")
