"""Build libwoxel_hip.so in-tree: HIP kernels for gfx950 + the C++ host builder.

    python woxel_amd/build.py            # by path: the package import needs the built library

Flags that the bit-exact contract depends on:
  -ffp-contract=off   no FMA contraction (host and device), matching the oracle
  no fast-math        IEEE-correct division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwoxel_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["trace.hip", "server.hip", "render.hip", "sdf_device.hip", "noise_device.hip", "grid_file.hip"]
CXX_SOURCES = ["vdb_host.cpp", "capi_host.cpp", "vdb_file.cpp"]
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra"]
# code object v5: loadable by the HIP runtime that PyTorch ships (ROCm 7.0) as well as ROCm 7.2
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-gpu-rdc", "-mcode-object-version=5"]


def _hipcc() -> str:
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    return r


def build(verbose: bool = False) -> str:
    """Compile every source and link libwoxel_hip.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run([_hipcc(), *HIP_FLAGS, *COMMON, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    cxx = shutil.which("g++") or "g++"
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run([cxx, *COMMON, "-pthread", "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
