"""Sedov blast against its analytic solution (physics-level validation, the
comparison src/sedov_blast/sedov_plot.py makes by eye in the reference).

The fixture tests/golden/sedov_analytic_2d.json holds the analytic profile of
F.X. Timmes' solver (tests/golden/make_sedov_analytic.py): cylindrical blast,
E = 0.311357, gamma = 1.4, uniform density 1, t = 0.1.  The numerical run is
the reference's test_sedov_blast_2d.ini (256 x 256, HLLC, MC slopes, CFL 0.8,
absorbing walls, blast radius 0.01 with the energy adjusted to the discrete
disk) stopped at t = 0.1, where the shock lies inside the box.  Convergence,
not parity: the tolerances are written below.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

FIXTURE = json.loads((Path(__file__).parent / "golden" / "sedov_analytic_2d.json").read_text())
NBINS = 200
BIN = np.hypot(0.5, 0.5) / NBINS   # the profile's bin width (half-diagonal / nbins)
L1_TOL = 0.12                      # relative L1 density error over r < 0.35 at 256^2
SHOCK_BINS = 2.0                   # density maximum within 2 bins of the shock radius


def sedov_params(n: int):
    return O.default_params(nx=n, ny=n, problemType=1, cfl=0.8, nStepmax=100000, tEnd=0.1,
                            boundary_type_xmin=2, boundary_type_xmax=2, boundary_type_ymin=2,
                            boundary_type_ymax=2, blast_density_in=1.0, blast_density_out=1.0,
                            blast_pressure_in=1.0, blast_pressure_out=1e-7, blast_radius=0.01,
                            blast_total_energy_inside=FIXTURE["params"]["eblast"],
                            slope_type=2.0, gamma0=1.4)


def compare(d, m):
    """Shock position, front position and relative L1 error of a numerical
    radial density profile (bin centres d, mean densities m) against the
    analytic one."""
    ana = np.interp(d, np.array(FIXTURE["r"]), np.array(FIXTURE["density"]))
    ok = np.isfinite(m)
    imax = int(np.nanargmax(m))
    half = (m[imax] + 1.0) / 2.0
    front = d[imax + int(np.argmax(m[imax:] < half))]
    sel = ok & (d < 0.35)
    l1 = float(np.abs(m[sel] - ana[sel]).sum() / ana[sel].sum())
    return float(d[imax]), float(front), l1


def test_fixture_closed_forms():
    """The analytic profile against what fixes it independently of the solver:
    the strong-shock density jump (gamma+1)/(gamma-1), the similarity radius
    R = (E t^2 / (alpha rho0))^(1/4) of a cylindrical blast with the energy
    integral alpha = 0.984074 for gamma = 1.4 (Kamm & Timmes 2007), the
    undisturbed density outside and the evacuated centre."""
    g = FIXTURE["params"]["gamma"]
    assert FIXTURE["rho2"] == pytest.approx((g + 1) / (g - 1), rel=1e-6)
    assert FIXTURE["alpha"] == pytest.approx(0.984074, abs=2e-6)
    e, t = FIXTURE["params"]["eblast"], FIXTURE["params"]["time"]
    assert FIXTURE["r2"] == pytest.approx((e * t * t / FIXTURE["alpha"]) ** 0.25, rel=1e-6)
    r, rho = np.array(FIXTURE["r"]), np.array(FIXTURE["density"])
    assert np.all(rho[r > FIXTURE["r2"] * 1.001] == 1.0)
    assert rho.max() <= FIXTURE["rho2"] and rho.max() > 0.95 * FIXTURE["rho2"]
    assert np.all(np.diff(rho[r < FIXTURE["r2"]]) >= 0)   # monotone behind the shock
    assert rho[0] < 1e-10


def oracle_profile(n: int):
    p = sedov_params(n)
    run = O.OracleRun(p)
    run.run()
    assert abs(run.t - 0.1) < 1e-15
    d, m, _ = O.radial_profile(run, run.nStep % 2, NBINS)
    return d, m


def test_oracle_matches_analytic_and_converges():
    d, m = oracle_profile(256)
    shock, front, l1 = compare(d, m)
    r2 = FIXTURE["r2"]
    print(f"256^2: density max at {shock:.4f}, half-max front {front:.4f} (r2 {r2:.4f}), "
          f"L1 {l1:.4f}")
    assert abs(shock - r2) <= SHOCK_BINS * BIN
    assert r2 - SHOCK_BINS * BIN <= front <= r2 + 3 * BIN
    assert l1 <= L1_TOL
    # halving the cell size brings the profile closer to the analytic one
    _, _, l1_coarse = compare(*oracle_profile(128))
    print(f"128^2: L1 {l1_coarse:.4f}")
    assert l1 < 0.8 * l1_coarse
