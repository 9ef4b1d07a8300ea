import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def hip_lib():
    from euler2d_kokkos_amd._capi import load_library
    return load_library()


@pytest.fixture(scope="session")
def gpu_available(hip_lib):
    from euler2d_kokkos_amd.hydro import device_info
    try:
        name, ncu = device_info()
    except Exception as e:  # noqa: BLE001
        pytest.fail(f"GPU test but no HIP device: {e}")
    print(f"[gpu] device {name} with {ncu} CUs", flush=True)
    return name


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a crash is attributable
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
        sys.stdout.flush()
    yield
