"""The C-ABI library loads on the CPU and exports every declared symbol."""
import ctypes as C
import re
from pathlib import Path

import pytest

from euler2d_kokkos_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(e2d_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_api():
    names = declared_symbols()
    assert "e2d_godunov_unsplit" in names and "e2d_compute_dt" in names
    assert len(names) >= 25


@pytest.mark.parametrize("name", declared_symbols())
def test_symbol_exported(hip_lib, name):
    assert hasattr(hip_lib, name), name


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_capi.EXPORTED) | {"e2d_version"}


def test_params_layout(hip_lib):
    assert hip_lib.e2d_params_size() == C.sizeof(_capi.E2DParams)
    import oracle_lib as O
    assert O.lib().ora_params_size() == C.sizeof(_capi.E2DParams)


def test_errors_without_device_do_not_abort(hip_lib, tmp_path):
    """Argument errors come back as status codes with a message."""
    assert hip_lib.e2d_create(None, 0, None) == _capi.E2D_ERR_ARG
    assert b"null" in hip_lib.e2d_last_error()
    p = _capi.E2DParams()
    # invalid params are rejected before any device work
    p.nx = 1
    assert hip_lib.e2d_params_validate(C.byref(p)) == _capi.E2D_ERR_UNSUPPORTED
    h = C.c_void_p()
    assert hip_lib.e2d_create(C.byref(p), 0, C.byref(h)) == _capi.E2D_ERR_UNSUPPORTED
    assert hip_lib.e2d_destroy(None) == _capi.E2D_ERR_ARG
    out = _capi.E2DParams()
    err = C.c_int(0)
    rc = hip_lib.e2d_params_from_ini(str(tmp_path / "missing.ini").encode(), C.byref(out),
                                     C.byref(err))
    assert rc == _capi.E2D_ERR_IO and err.value == -1
    # a missing deck still yields the reference's defaults (tEnd=0 -> no step)
    assert out.tEnd == 0.0 and out.nx == 2 and out.nStepmax == 1000
