""".ini parsing: the product's ConfigMap / HydroParams and the oracle's against
the reference's own reader (config/ConfigMap.cpp, config/inih/*), both through
the committed vectors of tests/golden/config_vectors.json and, when
oracle/_ref/libref_config.so is built, live."""
import ctypes as C
import json
import math
import struct
from pathlib import Path

import pytest

import oracle_lib as O
from config_queries import QUERIES
from euler2d_kokkos_amd import ConfigMap, HydroParams
from euler2d_kokkos_amd._capi import E2DParams

HERE = Path(__file__).resolve().parent
CASES = HERE / "golden" / "ini_cases"
VECTORS = json.loads((HERE / "golden" / "config_vectors.json").read_text())
DECKS = sorted((HERE.parent / "euler2d_kokkos_amd" / "decks").glob("*.ini"))


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def product_answer(cfg: ConfigMap, kind, sec, name, default):
    if kind == "string":
        return cfg.getString(sec, name, default)
    if kind == "integer":
        return cfg.getInteger(sec, name, default)
    if kind == "float":
        return f32_bits(cfg.getFloat(sec, name, default))
    return int(cfg.getBool(sec, name, bool(default)))


def oracle_answer(h, kind, sec, name, default):
    L = O.lib()
    s, n = sec.encode(), name.encode()
    if kind == "string":
        buf = C.create_string_buffer(512)
        L.ora_config_get_string(h, s, n, default.encode(), buf, 512)
        return buf.value.decode("latin-1")
    if kind == "integer":
        return L.ora_config_get_integer(h, s, n, default)
    if kind == "float":
        return f32_bits(L.ora_config_get_float(h, s, n, default))
    return L.ora_config_get_bool(h, s, n, int(default))


@pytest.mark.parametrize("case", sorted(VECTORS))
def test_product_parser_matches_reference_vectors(hip_lib, case):
    rec = VECTORS[case]
    cfg = ConfigMap(CASES / case)
    assert cfg.ParseError() == rec["parse_error"]
    for kind, sec, name, default, expected in rec["values"]:
        got = product_answer(cfg, kind, sec, name, default)
        assert got == expected, (case, kind, sec, name, got, expected)


@pytest.mark.parametrize("case", sorted(VECTORS))
def test_oracle_parser_matches_reference_vectors(case):
    rec = VECTORS[case]
    err = C.c_int(0)
    h = O.lib().ora_config_open(str(CASES / case).encode(), C.byref(err))
    try:
        assert err.value == rec["parse_error"]
        for kind, sec, name, default, expected in rec["values"]:
            got = oracle_answer(h, kind, sec, name, default)
            assert got == expected, (case, kind, sec, name, got, expected)
    finally:
        O.lib().ora_config_close(h)


def test_vectors_match_live_reference():
    """Re-derive the committed vectors from the reference reader when it is built."""
    L = O.ref_config_lib()
    if L is None:
        pytest.skip("oracle/_ref/libref_config.so not built (reference tree absent)")
    for case, rec in VECTORS.items():
        h = L.refcfg_open(str(CASES / case).encode())
        try:
            assert L.refcfg_parse_error(h) == rec["parse_error"]
            for kind, sec, name, default, expected in rec["values"]:
                s, n = sec.encode(), name.encode()
                if kind == "string":
                    buf = C.create_string_buffer(512)
                    L.refcfg_get_string(h, s, n, default.encode(), buf, 512)
                    got = buf.value.decode("latin-1")
                elif kind == "integer":
                    got = L.refcfg_get_integer(h, s, n, default)
                elif kind == "float":
                    got = f32_bits(L.refcfg_get_float(h, s, n, default))
                else:
                    got = L.refcfg_get_bool(h, s, n, int(default))
                assert got == expected, (case, name)
        finally:
            L.refcfg_close(h)


def test_missing_file():
    cfg = ConfigMap("/nonexistent/deck.ini")
    assert cfg.ParseError() == -1
    assert cfg.getInteger("run", "nstepmax", 1000) == 1000
    err = C.c_int(0)
    h = O.lib().ora_config_open(b"/nonexistent/deck.ini", C.byref(err))
    assert err.value == -1
    O.lib().ora_config_close(h)


def _same(a: E2DParams, b: E2DParams):
    for name, _ in a._fields_:
        x, y = getattr(a, name), getattr(b, name)
        if isinstance(x, float) and math.isnan(x):
            assert math.isnan(y), name
        else:
            assert x == y, (name, x, y)


@pytest.mark.parametrize("deck", [*sorted(CASES.glob("*.ini")), *DECKS], ids=lambda p: p.name)
def test_hydroparams_product_equals_oracle(hip_lib, deck):
    prod = HydroParams.from_ini(deck).raw
    ora = O.params_from_ini(deck)
    _same(prod, ora)


def test_float_rounding_of_reals(hip_lib, tmp_path):
    """Reals are parsed with strtof then widened (ConfigMap.cpp:38)."""
    f = tmp_path / "d.ini"
    f.write_text("[hydro]\ngamma0=1.4\ncfl=0.8\nsmallc=1e-10\n[mesh]\nxmax=0.1\n")
    p = HydroParams.from_ini(f)
    assert p.gamma0 == 1.399999976158142 and p.gamma0 != 1.4
    assert p.cfl == 0.800000011920929
    assert p.smallc == 1.000000013351432e-10
    assert p.xmax == struct.unpack("<f", struct.pack("<f", 0.1))[0]
    # defaults pass through the float-typed default argument as well
    q = HydroParams.from_ini(tmp_path / "absent.ini")
    assert q.gamma0 == 1.399999976158142 and q.cfl == 0.5 and q.tEnd == 0.0
