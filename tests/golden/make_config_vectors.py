"""Writes tests/golden/config_vectors.json: what the reference's own .ini reader
(config/ConfigMap.cpp, config/inih/INIReader.cpp, config/inih/ini.cpp, compiled
unmodified into oracle/_ref/libref_config.so by oracle/Makefile) returns for
the decks in tests/golden/ini_cases/.

Run in a container where /root/reference exists:
    make -C oracle ref && python tests/golden/make_config_vectors.py
"""
import json
import struct
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

from config_queries import QUERIES  # noqa: E402
import oracle_lib as O  # noqa: E402


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def main() -> None:
    import ctypes as C

    L = O.ref_config_lib()
    if L is None:
        raise SystemExit("oracle/_ref/libref_config.so not built (needs the reference tree)")
    out = {}
    for ini in sorted((HERE / "ini_cases").glob("*.ini")):
        h = L.refcfg_open(str(ini).encode())
        rec = {"parse_error": L.refcfg_parse_error(h), "values": []}
        for kind, sec, name, default in QUERIES:
            s, n = sec.encode(), name.encode()
            if kind == "string":
                buf = C.create_string_buffer(512)
                L.refcfg_get_string(h, s, n, default.encode(), buf, 512)
                val = buf.value.decode("latin-1")
            elif kind == "integer":
                val = L.refcfg_get_integer(h, s, n, default)
            elif kind == "float":
                val = f32_bits(L.refcfg_get_float(h, s, n, default))
            else:
                val = L.refcfg_get_bool(h, s, n, int(default))
            rec["values"].append([kind, sec, name, default, val])
        L.refcfg_close(h)
        out[ini.name] = rec
    (HERE / "config_vectors.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print(f"wrote {len(out)} decks")


if __name__ == "__main__":
    main()
