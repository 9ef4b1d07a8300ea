"""The host-side C++ of the product (the .ini reader, HydroParams, the output
writers) built with AddressSanitizer + UndefinedBehaviorSanitizer and run on
the .ini corner-case decks and the shipped decks: any out-of-bounds access,
leak or undefined behaviour fails the test.  CPU only; needs g++ with the
sanitizer runtimes."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "euler2d_kokkos_amd" / "csrc"
DECKS = sorted((ROOT / "tests" / "golden" / "ini_cases").glob("*.ini")) + \
    sorted((ROOT / "euler2d_kokkos_amd" / "decks").glob("*.ini"))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("san") / "sanitize_driver"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           f"-I{ROOT / 'include'}", f"-I{CSRC}",
           str(ROOT / "tests" / "host" / "sanitize_driver.cpp"),
           str(CSRC / "config.cpp"), str(CSRC / "params.cpp"), str(CSRC / "io.cpp"),
           "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0 and "asan" in (res.stderr or "").lower():
        pytest.skip("sanitizer runtime not available: " + res.stderr[-300:])
    assert res.returncode == 0, res.stderr
    return out


def test_host_code_under_sanitizers(driver, tmp_path):
    assert DECKS
    res = subprocess.run([str(driver), str(tmp_path), *map(str, DECKS),
                          str(tmp_path / "missing.ini")],
                         capture_output=True, text=True, timeout=600,
                         env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1",
                              "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
                              "PATH": "/usr/bin:/bin"})
    assert res.returncode == 0, (res.stdout[-2000:], res.stderr[-4000:])
    assert f"checked {len(DECKS) + 1} decks" in res.stdout
    assert list(tmp_path.glob("*.vti")) and list(tmp_path.glob("*.npy"))  # the writers ran
