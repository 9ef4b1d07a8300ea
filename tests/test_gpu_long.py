"""The gate configurations at their own step counts.

FP64, the drop-in, against the CPU oracle: the implode 256x256 deck (500
steps) and the blast 1024x1536 deck (2000 steps) at full size, the
four-quadrant deck (HLL, 1000 steps) and the shocked-bubble deck (HLLC, 3000
steps) on reduced grids.  Gate: relative L1 <= 1e-12 per conserved variable;
the FP64 kernels reproduce the oracle's operations, so the full arrays (ghost
cells included) and t must also be equal bit for bit.

Split precision against the FP64 GPU run, at full grid size and full step
count (blast, four-quadrant, shocked bubble) and on the implode deck: relative
L1 density <= 1e-4, and on the closed (reflecting) domains total mass and
energy conserved to 1e-12.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import (E2D_LAYOUT_INTERIOR, E2D_PRECISION_FP32_SPLIT,
                                E2D_PRECISION_FP64, E2D_U, E2D_U2, HydroParams, HydroRun)

pytestmark = pytest.mark.gpu

DECKS = Path(__file__).resolve().parents[1] / "euler2d_kokkos_amd" / "decks"
RTOL_FP64 = 1e-12
L1_SPLIT = 1e-4
DRIFT = 1e-12


def deck(name, **over):
    p = O.params_from_ini(DECKS / name)
    for k, v in over.items():
        setattr(p, k, v)
    O.lib().ora_params_init(p)
    return p


def _fp64_against_oracle(p, tag):
    ora = O.OracleRun(p)
    ora.run()
    run = HydroRun(HydroParams(p.copy()))
    run.run()
    assert run.nStep == ora.nStep == p.nStepmax, (tag, run.nStep, ora.nStep)
    g = run.get_state(run.current())
    c = ora.current()
    err = O.rel_l1(O.interior(g), O.interior(c))
    print(f"{tag}: steps={run.nStep} t={run.t!r} relL1={err.tolist()} "
          f"bitwise={np.array_equal(g, c)}", flush=True)
    assert np.all(err <= RTOL_FP64), err
    assert run.t == ora.t, (run.t, ora.t)
    for which in (E2D_U, E2D_U2):
        np.testing.assert_array_equal(run.get_state(which), ora.get(which), err_msg=tag)


def test_implode_deck_500_steps_fp64(gpu_available):
    """decks/implode_256.ini as written: 256x256, HLLC, minmod, CFL 0.8, 500 steps."""
    p = deck("implode_256.ini")
    assert (p.nx, p.ny, p.nStepmax, p.problemType) == (256, 256, 500, 0)
    _fp64_against_oracle(p, "implode 256x256 x 500")


def test_blast_deck_2000_steps_fp64(gpu_available):
    """decks/blast_1024x1536.ini as written: 1024x1536, HLLC, CFL 0.7, 2000 steps."""
    p = deck("blast_1024x1536.ini")
    assert (p.nx, p.ny, p.nStepmax, p.problemType) == (1024, 1536, 2000, 1)
    _fp64_against_oracle(p, "blast 1024x1536 x 2000")


def test_four_quadrant_1000_steps_fp64(gpu_available):
    """The four-quadrant deck (HLL, transmissive, 1000 steps) on 512x512."""
    p = deck("four_quadrant_8192.ini", nx=512, ny=512)
    assert (p.nStepmax, p.riemannSolverType, p.honor_riemann) == (1000, 1, 1)
    _fp64_against_oracle(p, "four-quadrant 512x512 x 1000")


def test_shocked_bubble_3000_steps_fp64(gpu_available):
    """The shocked-bubble deck (HLLC, inflow/outflow in x, walls in y, 3000
    steps) on 768x384."""
    p = deck("shocked_bubble_16384x8192.ini", nx=768, ny=384)
    assert p.nStepmax == 3000
    _fp64_against_oracle(p, "shocked bubble 768x384 x 3000")


# ------------------------------------------------------------ split precision
def _l1_planes(run_a, run_b):
    """Per-variable relative L1 of two runs' current interiors, one variable
    at a time (the full grids hold up to 134 M cells)."""
    a = run_a.get_state(run_a.current(), E2D_LAYOUT_INTERIOR)
    b = run_b.get_state(run_b.current(), E2D_LAYOUT_INTERIOR)
    err = np.empty(4)
    for v in range(4):
        num = np.abs(a[v] - b[v]).sum()
        den = np.abs(b[v]).sum()
        err[v] = num / den if den > 0 else num
    return err


SPLIT_DECKS = {
    "implode_256": ("implode_256.ini", True),
    "blast_1024x1536": ("blast_1024x1536.ini", True),
    "four_quadrant_8192": ("four_quadrant_8192.ini", False),
    "shocked_bubble_16384x8192": ("shocked_bubble_16384x8192.ini", False),
}


@pytest.mark.parametrize("name", list(SPLIT_DECKS))
def test_split_full_length(gpu_available, name):
    fname, closed = SPLIT_DECKS[name]
    p = deck(fname)
    r64 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP64)
    r64.run()
    r32 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)
    s0 = r32.conserved_sums(E2D_U)
    r32.run()
    assert r32.nStep == r64.nStep == p.nStepmax
    s1 = r32.conserved_sums(r32.current())
    drift = np.abs(s1 - s0) / np.abs(s0).clip(1e-300)
    kin = r32.kinetic_cells(r32.current())
    err = _l1_planes(r32, r64)
    print(f"{name}: {p.nx}x{p.ny} steps={r32.nStep} t32={r32.t!r} t64={r64.t!r} "
          f"relL1={err.tolist()} drift={drift.tolist()} kinetic_cells={kin}", flush=True)
    assert err[0] <= L1_SPLIT, err
    assert abs(r32.t - r64.t) <= 1e-4 * r64.t
    if closed:
        assert drift[0] <= DRIFT and drift[1] <= DRIFT, drift
    r32.close()
    r64.close()
