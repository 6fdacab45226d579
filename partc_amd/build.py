"""Build recipe of the engine library (partc_amd/libpartc.so) and the oracle.

The library is compiled for gfx950 with hipcc directly: every .hip and .cpp
file under partc_amd/csrc becomes one object, linked into one shared library
that exports the C ABI of include/partc.h.  Objects go to partc_amd/_build.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpartc.so"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("PTC_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # keep the value path's multiply and add separately rounded, as the
    # reference computes them (fksCuckooTC.hpp:313,318)
    "-ffp-contract=off",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-variable",
    "-Wno-unused-function",
    "-Wno-unused-value",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [_hipcc(), *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [_hipcc(), *[f for f in COMMON if not f.startswith("--offload-arch")], "-x", "c++", "-c",
               str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(f"[{src.name}]\n{warn}", file=sys.stderr)
    return obj


def build_lib(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CLI_SRC = PKG / "cli" / "partc_tc.cpp"
CLI = PKG / "partc-tc"


def build_cli(verbose: bool = True) -> Path:
    """The command-line driver (the reference's main.cpp interface), a host
    program over the C ABI, linked against libpartc.so."""
    cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(CLI_SRC), f"-L{PKG}", "-lpartc",
           "-Wl,-rpath,$ORIGIN", "-o", str(CLI)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"cli build failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {CLI}")
    return CLI


def build_oracle(verbose: bool = True) -> None:
    """The checker: oracle/liboracle.so and, where the reference tree is
    present, oracle/_ref/libcxsparse.so."""
    res = subprocess.run(["make", "-s", "-C", str(ORACLE), "-B"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    if verbose and res.stdout.strip():
        print(res.stdout.strip())


if __name__ == "__main__":
    build_lib()
    build_cli()
    build_oracle()
