"""Build recipe for the engine's shared library (hipcc, gfx950).

``python -m cuco_amd.build`` (or ``cuco_amd.build.build()``) compiles every
HIP source under ``cuco_amd/csrc`` for gfx950 and links
``cuco_amd/libcuco_engine.so`` in-tree.  It always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuco_engine.so")
DRIVER = os.path.join(HERE, "cuco_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
    "-I", CSRC,
    "-Wall",
    "-Wno-unused-result",
]

# FP64 parity needs every multiply and add rounded on its own, exactly as the
# C reference writes them; only the FP32 pair kernel and the FP32 register
# FFTs may fuse.
CONTRACT = {"pp32.hip": "-ffp-contract=fast", "fft32.hip": "-ffp-contract=fast"}
SOURCES = ["engine.hip", "sort.hip", "pp64.hip", "pp32.hip", "pm.hip", "fft32.hip",
           "integrate.hip", "driftfac.cpp", "snapshot.cpp"]
CXX = os.environ.get("CXX", "g++")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):  # host-only C++
        cmd = [CXX, "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
               "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, CONTRACT.get(src, "-ffp-contract=off"), "-c",
               os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    # The C host program in the reference's shape, linked against the library.
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=c99", "-Wall", "-I", INCLUDE,
           os.path.join(HERE, "host", "driver.c"), "-o", DRIVER, "-L", HERE,
           "-lcuco_engine", "-Wl,-rpath,$ORIGIN", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{res.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD_DIR, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
