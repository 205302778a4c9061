"""Build libsdtfe_amd.so for gfx950 with hipcc (in-tree, no torch).

    python -m sdtfe_amd.build

Every translation unit is compiled with -ffp-contract=off, so the FP64
arithmetic keeps the reference's rounding (the mixed walk writes its FP32
FMAs explicitly), and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libsdtfe_amd.so")
BUILD_DIR = os.path.join(ROOT, "_build")

ARCH = os.environ.get("SDTFE_ARCH", "gfx950")
COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
          "-Wno-unused-command-line-argument"]
# translation unit -> extra flags
UNITS = {
    "field.hip": ["-ffp-contract=off"],
    "walk_fp64.hip": ["-ffp-contract=off"],
    "walk_mixed.hip": ["-ffp-contract=off"],
    "batch.hip": ["-ffp-contract=off"],
    "volume.hip": ["-ffp-contract=off"],
    "adapter.hip": ["-ffp-contract=off"],
    "utils_io.cpp": ["-ffp-contract=off"],
    "capi.cpp": ["-x", "hip", "-ffp-contract=off"],
    "server.cpp": ["-x", "hip", "-ffp-contract=off"],
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(unit: str) -> str:
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD_DIR, unit + ".o")
    cmd = [hipcc(), *COMMON, *UNITS[unit], "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(UNITS))) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
