"""FP64 HIP path against the CPU oracle on the same seeded inputs.

Gate of the FP64 drop-in: relative L1 <= 1e-12 per conserved variable after N
steps.  The FP64 kernels are built to reproduce the oracle bit for bit, so the
tests also check exact equality of the full arrays (ghost cells included) and
of the accumulated time t.
"""
import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import (E2D_LAYOUT_INTERIOR, E2D_LAYOUT_LEFT, E2D_SCHEDULE_ATOMIC,
                                E2D_SCHEDULE_DIRECTIONAL, E2D_SCHEDULE_FUSED,
                                E2D_SCHEDULE_PASSES, E2D_U, E2D_U2, HydroParams, HydroRun)

pytestmark = pytest.mark.gpu

RTOL_L1 = 1e-12
IU, IV = 2, 3  # momentum planes of the [i][j][var] arrays (ID 0, IP 1: the energy)

CASES = {
    "implode64": dict(nx=64, ny=64, problemType=0, cfl=0.8, nStepmax=60, tEnd=1e9),
    "implode_aniso": dict(nx=70, ny=45, problemType=0, cfl=0.8, nStepmax=40, tEnd=1e9,
                          xmin=-1.0, xmax=1.0, ymax=1.0, slope_type=2.0, gamma0=1.666),
    "blast48x72": dict(nx=48, ny=72, problemType=1, cfl=0.7, nStepmax=60, tEnd=1e9, ymax=1.5,
                       blast_radius=0.1, blast_center_x=0.5, blast_center_y=0.75,
                       blast_density_in=1.0, blast_density_out=1.0, blast_pressure_in=10.0,
                       blast_pressure_out=0.1),
    "four_quadrant64_bc2": dict(nx=64, ny=64, problemType=2, cfl=0.8, nStepmax=50, tEnd=1e9,
                                boundary_type_xmin=2, boundary_type_xmax=2,
                                boundary_type_ymin=2, boundary_type_ymax=2),
    "shocked_bubble96x48": dict(nx=96, ny=48, problemType=4, cfl=0.8, nStepmax=40, tEnd=1e9,
                                xmax=2.0, boundary_type_xmin=2, boundary_type_xmax=2,
                                sb_preshock_density=1.0, sb_preshock_pressure=1.0,
                                sb_postshock_density=3.857142857142857,
                                sb_postshock_pressure=10.333333333333332,
                                sb_postshock_velocity=2.6293687924887177, sb_shock_loc=0.1,
                                sb_bubble_density=0.1, sb_bubble_pressure=1.0,
                                sb_bubble_radius=0.15, sb_bubble_center_x=0.4,
                                sb_bubble_center_y=0.5),
    "discontinuity64": dict(nx=64, ny=64, problemType=3, cfl=0.8, nStepmax=50, tEnd=1e9),
    "periodic_blast": dict(nx=61, ny=67, problemType=1, cfl=0.6, nStepmax=40, tEnd=1e9,
                           boundary_type_xmin=3, boundary_type_xmax=3, boundary_type_ymin=3,
                           boundary_type_ymax=3, blast_radius=0.2, blast_center_x=0.3,
                           blast_center_y=0.6),
    "mixed_bc_slope0": dict(nx=33, ny=129, problemType=1, cfl=0.5, nStepmax=30, tEnd=1e9,
                            boundary_type_xmin=1, boundary_type_xmax=2, boundary_type_ymin=3,
                            boundary_type_ymax=3, slope_type=0.0, blast_radius=0.25),
    "tend_clamp": dict(nx=40, ny=40, problemType=0, cfl=0.8, nStepmax=1000, tEnd=0.05),
    "four_quadrant_hll": dict(nx=64, ny=48, problemType=2, cfl=0.8, nStepmax=40, tEnd=1e9,
                              boundary_type_xmin=2, boundary_type_xmax=2, boundary_type_ymin=2,
                              boundary_type_ymax=2, honor_riemann=1, riemannSolverType=1),
    "blast_approx": dict(nx=48, ny=56, problemType=1, cfl=0.6, nStepmax=40, tEnd=1e9,
                         honor_riemann=1, riemannSolverType=0, blast_radius=0.15),
    "implode_approx_reflect": dict(nx=40, ny=40, problemType=0, cfl=0.7, nStepmax=40,
                                   tEnd=1e9, honor_riemann=1, riemannSolverType=0),
    "sedov_energy": dict(nx=50, ny=50, problemType=1, cfl=0.7, nStepmax=30, tEnd=1e9,
                         blast_total_energy_inside=1.0, blast_radius=0.05,
                         blast_density_out=1.0),
    # smallc^2 below 2^-759: the wave speeds take the full square root
    "blast_tiny_smallc": dict(nx=40, ny=48, problemType=1, cfl=0.7, nStepmax=30, tEnd=1e9,
                              smallc=1e-200, blast_radius=0.15),
    # smallr * smallp below 2^-1000: the sound speed of the CFL rate keeps the full root
    "implode_tiny_smallr": dict(nx=36, ny=44, problemType=0, cfl=0.8, nStepmax=30, tEnd=1e9,
                                smallr=1e-300),
}


def _params(case):
    return O.default_params(**CASES[case])


def _gpu_run(p, schedule, steps=0):
    hp = HydroParams(p.copy())
    run = HydroRun(hp, schedule=schedule)
    run.run(steps)
    return run


def _compare(run, ora, tag):
    assert run.nStep == ora.nStep, (tag, run.nStep, ora.nStep)
    assert run.t == ora.t, (tag, run.t, ora.t)
    g = run.get_state(run.current())
    c = ora.current()
    err = O.rel_l1(O.interior(g), O.interior(c))
    print(f"{tag}: steps={run.nStep} t={run.t!r} relL1={err} bitwise={np.array_equal(g, c)}")
    assert np.all(err <= RTOL_L1), err
    # both arrays, ghosts included, equal the oracle's
    for which in (E2D_U, E2D_U2):
        np.testing.assert_array_equal(run.get_state(which), ora.get(which), err_msg=tag)


@pytest.mark.parametrize("schedule", [E2D_SCHEDULE_FUSED, E2D_SCHEDULE_PASSES,
                                      E2D_SCHEDULE_DIRECTIONAL],
                         ids=["fused", "passes", "directional"])
@pytest.mark.parametrize("case", list(CASES))
def test_loop_matches_oracle(gpu_available, case, schedule):
    p = _params(case)
    ora = O.OracleRun(p)
    ora.run()
    run = _gpu_run(p, schedule)
    _compare(run, ora, f"{case}/{schedule}")


@pytest.mark.parametrize("case", list(CASES))
def test_loop_wide_addressing_matches_oracle(gpu_available, case, monkeypatch):
    """The run loop with the 64-bit row addressing that the step kernels use
    for planes above 4 GiB (forced here on small grids)."""
    monkeypatch.setenv("E2D_WIDE_ADDRESSING", "1")
    p = _params(case)
    ora = O.OracleRun(p)
    ora.run()
    run = _gpu_run(p, E2D_SCHEDULE_FUSED)
    _compare(run, ora, f"{case}/wide")


@pytest.mark.parametrize("tile", [3, 16, 28])
@pytest.mark.parametrize("bc", [1, 3], ids=["reflecting", "periodic"])
def test_interior_wavefronts_match_oracle(gpu_available, tile, bc):
    """Grids wide and tall enough for interior wavefronts (every loaded cell
    inside the domain: columns 60 s .. 60 s + 63 and rows tile k .. tile (k+1)
    + 3), which run a march without ghost handling, next to edge wavefronts."""
    p = O.default_params(nx=250, ny=181, problemType=1, cfl=0.7, nStepmax=25, tEnd=1e9,
                         blast_radius=0.2, blast_center_x=0.45, blast_center_y=0.6,
                         boundary_type_xmin=bc, boundary_type_xmax=bc,
                         boundary_type_ymin=bc, boundary_type_ymax=bc)
    ora = O.OracleRun(p)
    ora.run()
    run = HydroRun(HydroParams(p.copy()))
    run.set_tile(tile)
    run.run()
    _compare(run, ora, f"interior/{tile}/{bc}")


def test_two_wavefronts_per_simd_matches_oracle(gpu_available):
    """The blast workload's own grid, 1024 x 1536, at the 14-row tile: 2016
    wavefronts, which on a 256-CU MI355X are resident two per SIMD at once, so
    each march also steps its issue priority down as it progresses."""
    p = O.default_params(nx=1024, ny=1536, ymax=1.5, problemType=1, cfl=0.7, nStepmax=6,
                         tEnd=1e9, blast_radius=0.1, blast_center_x=0.5, blast_center_y=0.75,
                         blast_density_in=1.0, blast_density_out=1.0, blast_pressure_in=10.0,
                         blast_pressure_out=0.1)
    ora = O.OracleRun(p)
    ora.run()
    run = HydroRun(HydroParams(p.copy()))
    run.set_tile(14)
    run.run()
    _compare(run, ora, "blast1024x1536/14")


def test_planes_above_4gib(gpu_available):
    """1000 x 540000 cells: each FP64 plane holds 4.35 GB, so the step kernel
    addresses rows with 64-bit offsets; the blast sits at rows ~539000, beyond
    the 4 GiB offset.  Two steps of the fused loop must equal two steps of the
    reference pass schedule (equal time and equal conserved sums, which the
    same deterministic reduction computes from both arrays)."""
    p = O.default_params(nx=1000, ny=540000, ymax=540.0, problemType=1, cfl=0.7,
                         nStepmax=2, tEnd=1e9, blast_radius=0.05, blast_center_x=0.5,
                         blast_center_y=539.0)
    res = []
    for schedule in (E2D_SCHEDULE_FUSED, E2D_SCHEDULE_PASSES):
        run = HydroRun(HydroParams(p.copy()), schedule=schedule)
        run.run()
        res.append((run.nStep, run.t, run.conserved_sums(run.current()).tolist()))
        print(schedule, res[-1], flush=True)
        run.close()
        del run
    assert res[0][0] == 2
    assert res[0] == res[1]


@pytest.mark.parametrize("case", ["implode64", "blast48x72"])
def test_reference_api_steps(gpu_available, case):
    """compute_dt / godunov_unsplit / make_boundaries called step by step."""
    p = _params(case)
    ora = O.OracleRun(p)
    hp = HydroParams(p.copy())
    run = HydroRun(hp)
    assert run.compute_dt(0) == ora.compute_dt(0)
    for which in (0, 1):
        ora.make_boundaries(which)
        run.make_boundaries(which)
        np.testing.assert_array_equal(run.get_state(which), ora.get(which))
    t = 0.0
    for n in range(12):
        dt_o = ora.compute_dt(n % 2)
        dt_g = run.compute_dt(n % 2)
        assert dt_g == dt_o, (n, dt_g, dt_o)
        ora.godunov_unsplit(n, dt_o)
        run.godunov_unsplit(n, dt_g)
        t += dt_o
    for which in (0, 1):
        np.testing.assert_array_equal(run.get_state(which), ora.get(which))


def test_layouts_roundtrip(gpu_available):
    p = _params("implode_aniso")
    run = HydroRun(HydroParams(p.copy()))
    rng = np.random.default_rng(7)
    a = rng.random((p.isize, p.jsize, 4)) + 0.5
    run.set_state(E2D_U, a)
    np.testing.assert_array_equal(run.get_state(E2D_U), a)
    left = run.get_state(E2D_U, E2D_LAYOUT_LEFT)
    np.testing.assert_array_equal(left, np.transpose(a, (2, 1, 0)))
    inter = run.get_state(E2D_U, E2D_LAYOUT_INTERIOR)
    np.testing.assert_array_equal(inter, np.transpose(a[2:-2, 2:-2, :], (2, 1, 0)))
    b = rng.random(inter.shape)
    run.set_state(E2D_U2, b, E2D_LAYOUT_INTERIOR)
    np.testing.assert_array_equal(run.get_state(E2D_U2, E2D_LAYOUT_INTERIOR), b)


def test_conserved_sums_reflecting(gpu_available):
    """Reflecting walls: mass and energy conserved to summation roundoff."""
    p = _params("blast48x72")
    run = HydroRun(HydroParams(p.copy()))
    s0 = run.conserved_sums(E2D_U)
    ref0 = O.interior(run.get_state(E2D_U)).reshape(-1, 4).sum(0)
    np.testing.assert_allclose(s0, ref0, rtol=1e-14)
    run.run(200)
    s1 = run.conserved_sums(run.current())
    assert abs(s1[0] - s0[0]) <= 1e-12 * abs(s0[0])
    assert abs(s1[1] - s0[1]) <= 1e-12 * abs(s0[1])


def test_loop_in_pieces_matches_oracle(gpu_available):
    """Run-loop calls of uneven lengths continue one loop (each call restarts
    the device-side loop control from the state the last one left), and the
    launches queued past tEnd change nothing: the last call asks for far more
    steps than remain."""
    p = _params("tend_clamp")
    ora = O.OracleRun(p)
    ora.run()
    run = HydroRun(HydroParams(p.copy()))
    for k in (2, 3, 1):
        run.run(k)
    assert run.nStep == 6 and ora.nStep == 9
    run.run(40)
    _compare(run, ora, "pieces")
    assert run.t == p.tEnd


def test_loop_graph_replays_match_oracle(gpu_available):
    """Run-loop calls of 30+ steps replay a captured graph of 30 launches where
    the control ring's phase is 0 and the step even.  Every replay is checked
    against the oracle at the same step, with eager single launches, state
    downloads and a tile change (a new capture) in between."""
    p = _params("implode64")
    p.nStepmax = 400
    ora = O.OracleRun(p)
    run = HydroRun(HydroParams(p.copy()))
    run.set_tile(9)
    for k, tile in ((30, None), (3, None), (33, None), (60, 16), (1, None), (61, None)):
        if tile:
            run.set_tile(tile)
        run.run(k)
        ora.run(k)
        _compare(run, ora, f"graph/{run.nStep}")


def test_autotune_restores_state(gpu_available):
    p = _params("blast48x72")
    run = HydroRun(HydroParams(p.copy()))
    run.run(5)
    before = [run.get_state(w) for w in (E2D_U, E2D_U2)]
    t, n = run.t, run.nStep
    best, ms = run.autotune(candidates=(4, 9, 16), steps=3)
    assert best in (4, 9, 16) and ms > 0 and run.get_tile() == best
    best, ms = run.autotune(steps=2)  # derived candidates
    assert 1 <= best <= 4096 and ms > 0 and run.get_tile() == best
    assert (run.t, run.nStep) == (t, n)
    for w, b in zip((E2D_U, E2D_U2), before):
        np.testing.assert_array_equal(run.get_state(w), b)
    # the tuned run continues exactly like an untuned one
    ora = O.OracleRun(p)
    ora.run(5)
    run.run()
    ora.run()
    _compare(run, ora, "autotuned")


@pytest.mark.parametrize("schedule", [E2D_SCHEDULE_FUSED, E2D_SCHEDULE_PASSES],
                         ids=["fused", "passes"])
def test_fp64_prim_quotient_fallback(gpu_available, schedule):
    """The FP64 primitive conversion divides the momenta and the energy by the
    density with one shared reciprocal, and a wavefront whose quotients leave
    the range where that equals IEEE division takes the IEEE division.  Here
    30% of the cells carry momenta of 0, -0, tiny, denormal and near-threshold
    magnitudes (both sides of 2^-969): every step's state and t still equal the
    oracle's bit for bit, so both the fast path and the fallback are exact."""
    p = _params("blast48x72")
    p.nStepmax = 8
    ora = O.OracleRun(p)
    a = ora.get(0)
    rng = np.random.default_rng(11)
    vals = np.array([0.0, -0.0, 1e-300, -1e-300, 5e-320, -5e-320, 1e-310, 2.0**-1000,
                     -(2.0**-969), 2.0**-968, 3e-292])
    mask = rng.random(a.shape[:2]) < 0.3
    for v in (IU, IV):
        a[..., v][mask] = vals[rng.integers(0, len(vals), size=int(mask.sum()))]
    for which in (0, 1):
        ora.set(which, a)
    run = HydroRun(HydroParams(p.copy()), schedule=schedule)
    run.set_state(E2D_U, a)
    run.set_state(E2D_U2, a)
    ora.run()
    run.run()
    _compare(run, ora, f"tiny momenta / {schedule}")


@pytest.mark.parametrize("schedule", [E2D_SCHEDULE_FUSED, E2D_SCHEDULE_PASSES],
                         ids=["fused", "passes"])
def test_fp64_source_quotient_fallback(gpu_available, schedule):
    """The fused step's trace divides the two pressure slopes by the density
    with the reciprocal kept from the primitive conversion, and one wave-level
    check per row covers those quotients and the conversion's.  Every third
    row holds, over 40 columns at rest, energies of 1e-280 down to 5e-320
    (both sides of 2^-969): the pressure slopes along x there are below the
    range of the shared form, while the rows above and below convert exactly,
    so the check fails on the source terms alone.  smallc = 1e-200 takes the
    pressure floor to zero (and the wave speeds to the full square root).  Every
    step's state and t equal the oracle's bit for bit."""
    p = O.default_params(**dict(CASES["blast48x72"], smallc=1e-200, nStepmax=8))
    assert p.smallp == 0.0
    ora = O.OracleRun(p)
    a = ora.get(0)
    rng = np.random.default_rng(5)
    vals = np.array([1e-280, 3e-292, 2.0**-960, 2.0**-968, 2.0**-970, 2.0**-1000, 1e-300,
                     1e-310, 5e-320])
    for j in range(2, a.shape[1] - 2, 3):
        a[5:45, j, 1] = vals[rng.integers(0, len(vals), size=40)]
        a[5:45, j, IU] = 0.0
        a[5:45, j, IV] = 0.0
    for which in (0, 1):
        ora.set(which, a)
    run = HydroRun(HydroParams(p.copy()), schedule=schedule)
    run.set_state(E2D_U, a)
    run.set_state(E2D_U2, a)
    ora.run()
    run.run()
    assert np.isfinite(ora.current()).all()
    _compare(run, ora, f"tiny pressure slopes / {schedule}")


def test_autotune_timing_matches_time_loop(gpu_available):
    """A candidate's ms per step from the autotuner (60-step runs, whose loop
    graphs are captured before the timed region) agrees with time_loop's on the
    same tile after a warm-up, within the spread of the engine clock."""
    p = O.default_params(nx=1024, ny=1536, problemType=1, cfl=0.7, nStepmax=10**6, tEnd=1e9,
                         ymax=1.5, blast_center_y=0.75)
    run = HydroRun(HydroParams(p.copy()))
    run.time_loop(200)  # the engine clock settles
    best, ms_tune = run.autotune(candidates=(14,), steps=60)
    assert best == 14
    run.time_loop(60)
    ms, n = run.time_loop(60)
    print(f"autotune {ms_tune * 1e3:.2f} us/step, time_loop {ms / n * 1e3:.2f} us/step",
          flush=True)
    assert n == 60 and 0.8 <= ms_tune / (ms / n) <= 1.25


@pytest.mark.parametrize("case", ["implode64", "blast48x72", "four_quadrant64_bc2",
                                  "periodic_blast", "four_quadrant_hll"])
def test_atomic_schedule_l1(gpu_available, case):
    """Implementation version 2 adds every face flux to both neighbours with
    FP64 atomics: the order of the additions, and so the last bits, are not
    fixed.  Gate: relative L1 <= 1e-12 per conserved variable."""
    p = _params(case)
    ora = O.OracleRun(p)
    ora.run()
    run = _gpu_run(p, E2D_SCHEDULE_ATOMIC)
    assert run.nStep == ora.nStep
    assert abs(run.t - ora.t) <= 1e-12 * abs(ora.t)
    err = O.rel_l1(O.interior(run.get_state(run.current())), O.interior(ora.current()))
    print(f"{case}/atomic relL1={err}")
    assert np.all(err <= RTOL_L1), err


@pytest.mark.parametrize("version,schedule", [(0, E2D_SCHEDULE_FUSED),
                                              (1, E2D_SCHEDULE_DIRECTIONAL),
                                              (2, E2D_SCHEDULE_ATOMIC)])
def test_implementation_version_selects_schedule(gpu_available, version, schedule):
    p = O.default_params(nx=32, ny=24, nStepmax=5, tEnd=1e9, implementationVersion=version)
    run = HydroRun(HydroParams(p.copy()))
    assert run.get_schedule() == schedule
    run.run()
    ora = O.OracleRun(p)
    ora.run()
    err = O.rel_l1(O.interior(run.get_state(run.current())), O.interior(ora.current()))
    assert np.all(err <= RTOL_L1), err
