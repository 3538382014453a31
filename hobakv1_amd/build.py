"""Build recipe for libhobak_hip.so (gfx950 only) and the C++ host facade test driver.

    python -m hobakv1_amd.build          # compile everything in-tree

The library is built in-tree so that the same .so travels with the repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhobak_hip.so")

HIP_SOURCES = ["element.hip", "solver_kernels.hip", "api.hip"]

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    # roctx ranges (rocprofiler-sdk's marker library) around the engine's phases
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


DRIVER = os.path.join(HERE, "host", "facade_driver")


def build_facade_driver() -> str:
    """The C++ façade's test driver (host-only code linked against libhobak_hip.so)."""
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "host", "facade_driver.cpp"), "-o", DRIVER + ".tmp",
          "-L", HERE, "-lhobak_hip", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def main():
    build_library()
    build_facade_driver()
    print("built", LIB, DRIVER)


if __name__ == "__main__":
    sys.exit(main())
