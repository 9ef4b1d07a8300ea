"""Builds libeuler2d_hip.so (and the euler2d CLI) in-tree with hipcc for gfx950.

Flags that matter for bit-exact FP64 results:
  -ffp-contract=off   no a*b+c -> fma contraction (host and device)
  no -ffast-math      correctly rounded IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions

Usage: python -m euler2d_kokkos_amd.build [--clean]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJDIR = PKG / "_build"
LIB = PKG / "libeuler2d_hip.so"
CLI = PKG / "euler2d"

ARCH = os.environ.get("E2D_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-command-line-argument",
]

# roctx ranges (rocprofv3 --marker-trace) named as the reference's profiling regions
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ROCTX = [f"-L{ROCM / 'lib'}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM / 'lib'}"]

SOURCES = ["config.cpp", "params.cpp", "api.cpp", "io.cpp", "kernels.hip", "noalloc.hip"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build")
    return exe


def _compile(src: str) -> Path:
    out = OBJDIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, "-c", str(CSRC / src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def build(verbose: bool = False) -> Path:
    """Compiles every HIP/C++ source and links the shared library and CLI."""
    OBJDIR.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), *ROCTX]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    cli_cmd = [hipcc(), *COMMON, str(CSRC / "main.cpp"), "-o", str(CLI), f"-L{PKG}",
               "-leuler2d_hip", "-Wl,-rpath,$ORIGIN", *ROCTX]
    res = subprocess.run(cli_cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cli_cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


def clean() -> None:
    for p in (LIB, CLI):
        if p.exists():
            p.unlink()
    if OBJDIR.exists():
        shutil.rmtree(OBJDIR)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    build(verbose=True)
