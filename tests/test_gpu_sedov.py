"""Sedov blast on the GPU against its analytic solution: the FP64 path and
the split-precision path, with the radial profile computed by the product's
kernel.  Same fixture, run and tolerances as tests/test_sedov_analytic.py
(the FP64 state equals the oracle's bit for bit; the split path is checked
here at the physics level)."""
import pytest

from euler2d_kokkos_amd import E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64, HydroParams, HydroRun
from test_sedov_analytic import BIN, FIXTURE, L1_TOL, NBINS, SHOCK_BINS, compare, sedov_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [E2D_PRECISION_FP64, E2D_PRECISION_FP32_SPLIT],
                         ids=["fp64", "split"])
def test_sedov_matches_analytic(gpu_available, precision):
    run = HydroRun(HydroParams(sedov_params(256)), precision=precision)
    run.run()
    assert abs(run.t - 0.1) < 1e-12
    d, m, _ = run.radial_profile(NBINS, which=run.current())
    shock, front, l1 = compare(d, m)
    r2 = FIXTURE["r2"]
    print(f"density max at {shock:.4f}, half-max front {front:.4f} (r2 {r2:.4f}), L1 {l1:.4f}")
    assert abs(shock - r2) <= SHOCK_BINS * BIN
    assert r2 - SHOCK_BINS * BIN <= front <= r2 + 3 * BIN
    assert l1 <= L1_TOL
