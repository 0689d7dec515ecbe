"""In-tree build of the RRSPH_HIP shared library for gfx950.

    python -m sph2d_amd.build        -> sph2d_amd/librrsph_hip.so

The library is the C-ABI of include/rrsph_hip.h; hipcc cross-compiles it
without a GPU.  Every HIP compile disables packed-FP32 instruction selection
(-packed-fp32-ops) so that float code is emitted as scalar VALU operations.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "rrsph_hip.hip")
OUT = os.path.join(HERE, "librrsph_hip.so")
HEADERS = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc")) if f.endswith(".hpp")]
HEADERS.append(os.path.join(ROOT, "include", "rrsph_hip.h"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(force: bool = True, verbose: bool = False) -> str:
    if not force and os.path.exists(OUT):
        newest = max(os.path.getmtime(p) for p in [SRC] + HEADERS)
        if os.path.getmtime(OUT) >= newest:
            return OUT
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp")
DRIVER_OUT = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


def build_dropin_driver() -> str:
    """Host-only C++ program that drives the library through the drop-in
    adapter include/rrsph_hip.hpp (g++, linked against librrsph_hip.so)."""
    tmp = DRIVER_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp, DRIVER_SRC,
           "-L", HERE, "-lrrsph_hip", "-Wl,-rpath,$ORIGIN/../../sph2d_amd", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed for " + DRIVER_SRC)
    os.replace(tmp, DRIVER_OUT)
    return DRIVER_OUT


if __name__ == "__main__":
    print(build(force=True, verbose="-v" in sys.argv))
    print(build_dropin_driver())
