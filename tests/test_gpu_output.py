"""On-disk output of GPU runs: HydroRun::saveVTK files, the Sedov radial
profile and the euler2d CLI's output cadence, against the CPU oracle (whose
states equal the GPU's bit for bit, so the .vti files must be identical)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import E2D_PRECISION_FP64, E2D_U, E2D_U2, HydroParams, HydroRun

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parents[1] / "euler2d_kokkos_amd" / "euler2d"


def test_save_vtk_equals_oracle(gpu_available, tmp_path):
    p = O.default_params(nx=40, ny=30, problemType=1, blast_radius=0.15, nStepmax=15, tEnd=1e9)
    run = HydroRun(HydroParams(p.copy()), precision=E2D_PRECISION_FP64)
    run.run()
    ora = O.OracleRun(p)
    ora.run()
    assert run.nStep == ora.nStep == 15
    for which in (E2D_U, E2D_U2):
        a = run.save_vtk(which, tmp_path, f"gpu{which}", 15)
        b = O.save_vtk(ora, which, tmp_path, f"ora{which}", 15)
        assert a.read_bytes() == b.read_bytes(), which


def test_radial_profile_equals_oracle(gpu_available):
    p = O.default_params(nx=64, ny=48, problemType=1, blast_radius=0.1,
                         blast_total_energy_inside=1.0, nStepmax=20, tEnd=1e9)
    run = HydroRun(HydroParams(p.copy()))
    run.run()
    ora = O.OracleRun(p)
    ora.run()
    for nbins in (37, 100, 1500):  # 1500 > the LDS bin limit of the kernel
        d, m, c = run.radial_profile(nbins, which=E2D_U)
        d0, m0, c0 = O.radial_profile(ora, 0, nbins)
        np.testing.assert_array_equal(c, c0)
        np.testing.assert_array_equal(d, d0)
        # FP64 sums in a different order: equal to rounding
        np.testing.assert_allclose(m, m0, rtol=1e-13, equal_nan=True)


def write_deck(path: Path, text: str) -> Path:
    path.write_text(text)
    return path


def test_cli_output_cadence(gpu_available, tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    deck = write_deck(tmp_path / "implode.ini", f"""
[run]
solver_name=Hydro_Muscl_2D
tEnd=10.0
nStepmax=25
nOutput=10
[mesh]
nx=32
ny=24
boundary_type_xmin=1
boundary_type_xmax=1
boundary_type_ymin=1
boundary_type_ymax=1
[hydro]
gamma0=1.4
cfl=0.8
riemann=hllc
problem=implode
[output]
outputDir={out}
outputPrefix=imp
""")
    res = subprocess.run([str(CLI), str(deck)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    files = sorted(x.name for x in out.iterdir())
    assert files == [f"imp_{k:07d}.vti" for k in (0, 10, 20, 25)]
    p = O.params_from_ini(deck)
    ora = O.OracleRun(p)
    for k in (0, 10, 20, 25):
        if k > ora.nStep:
            ora.run(k - ora.nStep)
        assert ora.nStep == k
        ref = O.save_vtk(ora, k % 2, tmp_path, "ora", k)
        assert (out / f"imp_{k:07d}.vti").read_bytes() == ref.read_bytes(), k


def test_cli_sedov_profile(gpu_available, tmp_path):
    deck = write_deck(tmp_path / "sedov.ini", """
[run]
tEnd=10.0
nStepmax=31
nOutput=-1
[mesh]
nx=48
ny=48
[hydro]
problem=blast
cfl=0.7
[blast]
radius=0.05
total_energy_inside=1.0
nbins=40
""")
    res = subprocess.run([str(CLI), str(deck)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert not list(tmp_path.glob("*.vti"))  # nOutput=-1 disables output
    d = np.load(tmp_path / "sedov_blast_radial_distances.npy")
    m = np.load(tmp_path / "sedov_blast_density_profile.npy")
    p = O.params_from_ini(deck)
    ora = O.OracleRun(p)
    ora.run()
    assert ora.nStep == 31
    d0, m0, _ = O.radial_profile(ora, 0, 40)  # the reference profiles U whatever the parity
    np.testing.assert_array_equal(d, d0)
    np.testing.assert_allclose(m, m0, rtol=1e-13, equal_nan=True)
