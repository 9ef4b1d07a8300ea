"""Pins of the oracle's HydroParams::init derived fields by what they do, not by
their formulas (HydroParams.cpp:160-176):

  smallp   the pressure floor per unit density: a state at the floor has the
           floor sound speed, sqrt(gamma0 * smallp) == smallc (to 1 ulp)
  smallpp  the pressure floor of a cell at the density floor: the primitive
           conversion of an empty cell returns it as the pressure
  dx, dy   (xmax - xmin) / nx with the float-parsed bounds: nx cells tile the
           interval (exactly on power-of-two grids)
  isize    nx + 4 (two ghost layers), the size of the arrays the oracle holds

gamma6 = (gamma0 + 1) / (2 gamma0) is pinned by the two-shock Riemann solution
(test_oracle_pointwise.py::test_approx_matches_exact_two_shock), where it sets
the shock mass flux.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

SMALLS = [(1e-10, 1e-10, 1.4), (1e-6, 3e-4, 1.666), (2.5e-3, 1e-8, 5.0 / 3.0),
          (1e-30, 1e-20, 1.1), (0.5, 0.25, 3.0)]


def _ulps(a, b):
    return abs(a - b) / math.ulp(b)


@pytest.mark.parametrize("smallr,smallc,gamma0", SMALLS)
def test_smallp_gives_floor_sound_speed(smallr, smallc, gamma0):
    P = O.default_params(smallr=smallr, smallc=smallc, gamma0=gamma0)
    # the sound speed of a state at the pressure floor is the floor sound speed
    assert _ulps(math.sqrt(P.gamma0 * P.smallp), P.smallc) <= 1.0, (P.smallp, P.smallc)
    # ... for any density: the floor is proportional to it
    for rho in (1e-3, 1.0, 7.5e2):
        _, c = O.prim(P, [rho, 0.0, 0.0, 0.0])  # no internal energy: p at the floor
        assert _ulps(c, P.smallc) <= 2.0, (rho, c, P.smallc)


@pytest.mark.parametrize("smallr,smallc,gamma0", SMALLS)
def test_smallpp_is_floor_pressure_of_empty_cell(smallr, smallc, gamma0):
    P = O.default_params(smallr=smallr, smallc=smallc, gamma0=gamma0)
    q, c = O.prim(P, [0.0, 0.0, 0.0, 0.0])
    assert q[0] == P.smallr  # density floor
    assert q[1] == P.smallpp  # pressure floor at the density floor
    assert _ulps(c, P.smallc) <= 2.0
    # an independent value: smallr * smallc^2 / gamma0 within rounding (the
    # fields are set here as doubles, not parsed from a deck)
    want = smallr * smallc ** 2 / gamma0
    assert _ulps(P.smallpp, want) <= 4.0, (P.smallpp, want)


@pytest.mark.parametrize("nx,ny", [(64, 32), (256, 256), (1024, 1536)])
def test_cell_sizes_and_array_sizes(nx, ny):
    P = O.default_params(nx=nx, ny=ny)
    # nx cells tile the unit interval: exactly for powers of two, to rounding else
    for n, h in ((nx, P.dx), (ny, P.dy)):
        assert abs(n * h - 1.0) <= (0.0 if n & (n - 1) == 0 else 2.0 ** -52), (n, h)
    assert (P.isize, P.jsize) == (nx + 4, ny + 4)
    run = O.OracleRun(P, init=False)
    assert run.get(0).shape == (nx + 4, ny + 4, 4)


def test_float_parsed_bounds_set_dx(tmp_path):
    """Reals of a deck are parsed as float (strtof): 0.1 becomes 0.1f, and dx
    follows from the float value, not from the decimal one."""
    p = O.params_from_text(tmp_path, "[mesh]\nnx=10\nny=10\nxmin=0.1\nxmax=0.7\n")
    x0, x1 = float(np.float32(0.1)), float(np.float32(0.7))
    assert p.xmin == x0 and p.xmax == x1
    assert p.dx == (x1 - x0) / 10
    assert p.dx != (0.7 - 0.1) / 10
