"""FP32 split-precision path against the FP64 path on the GPU.

Gates of the split path: relative L1 density error <= 1e-4 against FP64 after N
steps, and total mass and energy conserved to ~1e-12 (FP64 sums of hi + lo).
"""
import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import (E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64, E2D_U, E2D_U2,
                                E2DError, HydroParams, HydroRun)
from euler2d_kokkos_amd import E2D_SCHEDULE_PASSES

pytestmark = pytest.mark.gpu

L1_TOL = 1e-4

CASES = {
    "implode128": dict(nx=128, ny=128, problemType=0, cfl=0.8, nStepmax=300, tEnd=1e9),
    "blast96x144": dict(nx=96, ny=144, problemType=1, cfl=0.7, nStepmax=300, tEnd=1e9, ymax=1.5,
                        blast_radius=0.1, blast_center_x=0.5, blast_center_y=0.75,
                        blast_density_in=1.0, blast_density_out=1.0),
    "four_quadrant_hll": dict(nx=128, ny=128, problemType=2, cfl=0.8, nStepmax=200, tEnd=1e9,
                              boundary_type_xmin=2, boundary_type_xmax=2, boundary_type_ymin=2,
                              boundary_type_ymax=2, honor_riemann=1, riemannSolverType=1),
    # the approx (two-shock) solver in FP32
    "blast_approx": dict(nx=96, ny=80, problemType=1, cfl=0.6, nStepmax=200, tEnd=1e9,
                         honor_riemann=1, riemannSolverType=0, blast_radius=0.15),
    "shocked_bubble": dict(nx=192, ny=96, problemType=4, cfl=0.8, nStepmax=200, tEnd=1e9,
                           xmax=2.0, boundary_type_xmin=2, boundary_type_xmax=2,
                           sb_preshock_density=1.0, sb_preshock_pressure=1.0,
                           sb_postshock_density=3.857142857142857,
                           sb_postshock_pressure=10.333333333333332,
                           sb_postshock_velocity=2.6293687924887177, sb_shock_loc=0.1,
                           sb_bubble_density=0.1, sb_bubble_pressure=1.0, sb_bubble_radius=0.15,
                           sb_bubble_center_x=0.4, sb_bubble_center_y=0.5),
}


def _runs(case):
    p = O.default_params(**CASES[case])
    r64 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP64)
    r32 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)
    return p, r64, r32


@pytest.mark.parametrize("case", list(CASES))
def test_split_close_to_fp64(gpu_available, case):
    p, r64, r32 = _runs(case)
    s0 = r32.conserved_sums(E2D_U)
    r64.run()
    r32.run()
    assert r32.nStep == r64.nStep == p.nStepmax
    a = r32.get_state(r32.current())
    b = r64.get_state(r64.current())
    err = O.rel_l1(O.interior(a), O.interior(b))
    drift = (r32.conserved_sums(r32.current()) - s0) / np.abs(s0).clip(1e-300)
    print(f"{case}: steps={r32.nStep} t32={r32.t!r} t64={r64.t!r} relL1={err.tolist()} "
          f"drift={drift.tolist()}")
    assert err[0] <= L1_TOL, err
    assert abs(r32.t - r64.t) <= 1e-4 * r64.t


def test_split_two_wavefronts_per_simd(gpu_available):
    """The blast workload's grid, 1024 x 1536, at the 14-row tile (2016
    wavefronts, two per SIMD on a 256-CU MI355X, with the progress-based issue
    priority): the split run stays within the L1 gate of the FP64 run, and on
    the closed domain conserves mass and energy."""
    p = O.default_params(nx=1024, ny=1536, ymax=1.5, problemType=1, cfl=0.7, nStepmax=100,
                         tEnd=1e9, blast_radius=0.1, blast_center_x=0.5, blast_center_y=0.75,
                         blast_density_in=1.0, blast_density_out=1.0, blast_pressure_in=10.0,
                         blast_pressure_out=0.1)
    runs = {}
    for prec in (E2D_PRECISION_FP64, E2D_PRECISION_FP32_SPLIT):
        r = HydroRun(HydroParams(p.copy()), precision=prec)
        r.set_tile(14)
        runs[prec] = r
    r64, r32 = runs[E2D_PRECISION_FP64], runs[E2D_PRECISION_FP32_SPLIT]
    s0 = r32.conserved_sums(E2D_U)
    r64.run()
    r32.run()
    assert r32.nStep == r64.nStep == 100
    err = O.rel_l1(O.interior(r32.get_state(r32.current())),
                   O.interior(r64.get_state(r64.current())))
    s1 = r32.conserved_sums(r32.current())
    print(f"1024x1536/14: relL1={err.tolist()} mass/energy drift={(s1 - s0)[[0, 1]].tolist()}")
    assert err[0] <= L1_TOL, err
    for v in (0, 1):
        assert abs(s1[v] - s0[v]) <= 1e-12 * abs(s0[v]), (v, s0, s1)


@pytest.mark.parametrize("bc,conserved", [(1, (0, 1)), (3, (0, 1, 2, 3))])
def test_split_conservation(gpu_available, bc, conserved):
    p = O.default_params(nx=160, ny=120, problemType=1, cfl=0.7, nStepmax=1000, tEnd=1e9,
                         blast_radius=0.12, blast_center_x=0.45, blast_center_y=0.55,
                         boundary_type_xmin=bc, boundary_type_xmax=bc, boundary_type_ymin=bc,
                         boundary_type_ymax=bc)
    r = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)
    s0 = r.conserved_sums(E2D_U)
    r.run()
    s1 = r.conserved_sums(r.current())
    # drift relative to the L1 norm of each variable (momentum starts at zero)
    scale = np.abs(r.get_state(r.current())[2:-2, 2:-2, :]).reshape(-1, 4).sum(0)
    rel = np.abs(s1 - s0) / scale
    print(f"bc={bc} steps={r.nStep} drift={rel.tolist()}")
    for v in conserved:
        assert rel[v] <= 1e-12, (v, rel)


def test_split_uniform_state(gpu_available):
    p = O.default_params(nx=70, ny=50, cfl=0.8, nStepmax=50, tEnd=1e9, boundary_type_xmin=3,
                         boundary_type_xmax=3, boundary_type_ymin=3, boundary_type_ymax=3)
    r = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT, initialise=False)
    w = np.array([1.3, 0.7 / 0.4 + 0.5 * 1.3 * (0.4 ** 2 + 0.25 ** 2), 1.3 * 0.4, -1.3 * 0.25])
    u = np.broadcast_to(w, (p.isize, p.jsize, 4)).copy()
    r.set_state(E2D_U, u)
    r.set_state(E2D_U2, u)
    r.run()
    got = r.get_state(r.current())[2:-2, 2:-2, :]
    np.testing.assert_allclose(got, np.broadcast_to(w, got.shape), rtol=1e-6)


def test_split_state_roundtrip(gpu_available):
    p = O.default_params(nx=33, ny=17)
    r = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)
    rng = np.random.default_rng(3)
    a = rng.random((p.isize, p.jsize, 4)) + 0.5
    r.set_state(E2D_U, a)
    b = r.get_state(E2D_U)
    # hi + lo holds 48 significant bits
    np.testing.assert_allclose(b, a, rtol=2.0 ** -46, atol=0)


def test_split_rejects_pass_schedule(gpu_available):
    p = O.default_params(nx=16, ny=16, nStepmax=2, tEnd=1e9)
    r = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT,
                 schedule=E2D_SCHEDULE_PASSES)
    with pytest.raises(E2DError):
        r.godunov_unsplit(0, 1e-3)


def test_split_tend_clamp(gpu_available):
    p = O.default_params(nx=40, ny=40, problemType=0, cfl=0.8, nStepmax=100000, tEnd=0.05)
    r64 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP64)
    r32 = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)
    r64.run()
    r32.run()
    assert r64.t == pytest.approx(0.05, rel=1e-15) and r32.t == pytest.approx(0.05, rel=1e-15)
    assert abs(r32.nStep - r64.nStep) <= 1


def test_plain_fp32_comparator(gpu_available):
    """The all-float comparator runs the same problem; the split state keeps the
    conserved sums far better than plain FP32 does."""
    from euler2d_kokkos_amd import E2D_PRECISION_FP32
    p = O.default_params(nx=160, ny=120, problemType=1, cfl=0.7, nStepmax=600, tEnd=1e9,
                         blast_radius=0.12, blast_center_x=0.45, blast_center_y=0.55,
                         boundary_type_xmin=3, boundary_type_xmax=3, boundary_type_ymin=3,
                         boundary_type_ymax=3)
    out = {}
    for name, prec in (("fp64", E2D_PRECISION_FP64), ("split", E2D_PRECISION_FP32_SPLIT),
                       ("fp32", E2D_PRECISION_FP32)):
        r = HydroRun(HydroParams(p.copy()), precision=prec)
        s0 = r.conserved_sums(E2D_U)
        r.run()
        s1 = r.conserved_sums(r.current())
        out[name] = (r.get_state(r.current()), np.abs(s1 - s0)[:2] / np.abs(s0[:2]))
    ref = O.interior(out["fp64"][0])
    l1_split = O.rel_l1(O.interior(out["split"][0]), ref)
    l1_fp32 = O.rel_l1(O.interior(out["fp32"][0]), ref)
    print(f"L1 split={l1_split.tolist()} fp32={l1_fp32.tolist()} "
          f"drift split={out['split'][1].tolist()} fp32={out['fp32'][1].tolist()}")
    assert l1_fp32[0] < 1e-2 and l1_split[0] <= L1_TOL
    assert np.all(out["split"][1] <= 1e-12)
    assert np.all(out["fp32"][1] > out["split"][1])


@pytest.mark.parametrize("precision", [E2D_PRECISION_FP32_SPLIT, E2D_PRECISION_FP64],
                         ids=["split", "fp64"])
def test_wide_addressing_same_bits(gpu_available, precision, monkeypatch):
    """The 64-bit row addressing (planes above 4 GiB; forced here) and the
    32-bit one give the same arrays bit for bit."""
    p = O.default_params(**CASES["shocked_bubble"])
    a = HydroRun(HydroParams(p.copy()), precision=precision)
    a.run(60)
    monkeypatch.setenv("E2D_WIDE_ADDRESSING", "1")
    b = HydroRun(HydroParams(p.copy()), precision=precision)
    b.run(60)
    assert (a.t, a.nStep) == (b.t, b.nStep)
    for w in (E2D_U, E2D_U2):
        np.testing.assert_array_equal(a.get_state(w), b.get_state(w))


def _hypersonic_state(p, mach, angle=0.3):
    """A density wave carried by a uniform flow at the given Mach number
    (p = 1, rho = 1 + 0.3 sin(2 pi x) cos(2 pi y)): an exact translation of the
    density field, with internal energy 1/(1 + gamma (gamma-1) M^2 / 2) of the
    total at rho = 1."""
    g = p.gamma0
    x = p.xmin + (np.arange(p.isize) - 1.5) * p.dx
    y = p.ymin + (np.arange(p.jsize) - 1.5) * p.dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    speed = mach * np.sqrt(g * 1.0 / 1.0)
    u, v = speed * np.cos(angle), speed * np.sin(angle)
    E = 1.0 / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, E, rho * u, rho * v], axis=-1)


def _pressure(a, gamma):
    r, E, mx, my = (a[..., k] for k in range(4))
    return (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my) / r)


def test_split_kinetic_fallback(gpu_available):
    """Mach 20 uniform flow carrying a density wave (periodic): the internal
    energy is about 0.9% of the total, so every cell falls below the default
    kin_frac = 0.1 and takes the FP64 primitive conversion.  The branch must
    fire (kin_frac = 0 gives other bits), keep the density within the L1 gate
    and bring the pressure closer to the FP64 run than the FP32 conversion."""
    base = dict(nx=128, ny=96, cfl=0.8, nStepmax=300, tEnd=1e9, boundary_type_xmin=3,
                boundary_type_xmax=3, boundary_type_ymin=3, boundary_type_ymax=3)
    out = {}
    for name, prec, kf in (("fp64", E2D_PRECISION_FP64, 0.1),
                           ("split", E2D_PRECISION_FP32_SPLIT, 0.1),
                           ("split_nofallback", E2D_PRECISION_FP32_SPLIT, 0.0)):
        p = O.default_params(kin_frac=kf, **base)
        r = HydroRun(HydroParams(p.copy()), precision=prec, initialise=False)
        u0 = _hypersonic_state(p, mach=20.0)
        r.set_state(E2D_U, u0)
        r.set_state(E2D_U2, u0)
        kin0 = r.kinetic_cells(E2D_U)
        r.run()
        assert r.nStep == p.nStepmax
        out[name] = (r.get_state(r.current()), kin0, r.t)
    ref = O.interior(out["fp64"][0])
    p64 = _pressure(ref, 1.4)
    res = {}
    for name in ("split", "split_nofallback"):
        a = O.interior(out[name][0])
        l1 = O.rel_l1(a, ref)
        lp = np.abs(_pressure(a, 1.4) - p64).sum() / np.abs(p64).sum()
        res[name] = (l1, lp)
        print(f"{name}: kinetic cells at start={out[name][1]} relL1={l1.tolist()} "
              f"pressure relL1={lp:.3e} t={out[name][2]!r}", flush=True)
    ncell = base["nx"] * base["ny"]
    assert out["split"][1] == ncell and out["split_nofallback"][1] == 0
    assert not np.array_equal(out["split"][0], out["split_nofallback"][0])
    for name in res:
        assert res[name][0][0] <= L1_TOL, (name, res[name])
    assert res["split"][1] < res["split_nofallback"][1], res


def test_kin_frac_parameter(gpu_available, tmp_path):
    """[hydro] kin_frac is read as a float like every real key; the default is
    0.1 and a negative value is refused."""
    from euler2d_kokkos_amd import HydroParams as HP
    f = tmp_path / "k.ini"
    f.write_text("[mesh]\nnx=8\nny=8\n[hydro]\nkin_frac=0.25\n")
    assert HP.from_ini(f).kin_frac == np.float32(0.25)
    f.write_text("[mesh]\nnx=8\nny=8\n")
    assert HP.from_ini(f).kin_frac == np.float32(0.1)
    p = O.default_params(nx=8, ny=8, kin_frac=-1.0)
    with pytest.raises(E2DError):
        HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP32_SPLIT)


@pytest.mark.parametrize("precision", ["split", "fp32"])
def test_tile_and_prefetch_distance_same_bits(gpu_available, precision):
    """Tiles of 3 and 14 rows (rows prefetched one ahead) and of 20, 28 and 48
    rows (two ahead) give the same bits, full arrays and t: each cell's
    arithmetic does not depend on which wavefront marches it or how far ahead
    it loads.  250 x 181 has interior and edge strips, bottom, middle and top
    segments."""
    from euler2d_kokkos_amd import E2D_PRECISION_FP32
    prec = E2D_PRECISION_FP32_SPLIT if precision == "split" else E2D_PRECISION_FP32
    p = O.default_params(nx=250, ny=181, problemType=1, cfl=0.7, nStepmax=60, tEnd=1e9,
                         blast_radius=0.2, blast_center_x=0.4, blast_center_y=0.6,
                         boundary_type_xmin=1, boundary_type_xmax=2, boundary_type_ymin=3,
                         boundary_type_ymax=3)
    ref = None
    for tile in (3, 14, 20, 28, 48):
        r = HydroRun(HydroParams(p.copy()), precision=prec)
        r.set_tile(tile)
        r.run()
        got = (r.t, r.nStep, r.get_state(E2D_U), r.get_state(E2D_U2))
        r.close()
        if ref is None:
            ref = got
            continue
        assert got[:2] == ref[:2], (tile, got[:2], ref[:2])
        np.testing.assert_array_equal(got[2], ref[2], err_msg=f"U tile {tile}")
        np.testing.assert_array_equal(got[3], ref[3], err_msg=f"U2 tile {tile}")
