"""Builds libclwilson.so (gfx950 HIP kernels + C ABI + host orchestration),
the clwilson command-line executable, and the host-compiled arithmetic check
library used by the CPU unit tests.

Usage: python -m clwilson_amd.build   (or clwilson_amd.build.build_all())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

LIB = os.path.join(PKG, "libclwilson.so")
EXE = os.path.join(PKG, "clwilson")
HOSTCHECK = os.path.join(PKG, "libclw_hostcheck.so")

ARCH = os.environ.get("CLW_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in the unified VGPR file: the AGPR form spills and
    # shuffles the 48 accumulator tiles of k_mfma_runs between iterations
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

LIB_SOURCES = ["clw_kernels.hip", "clw_runtime.hip", "cl_wilson.hip", "clw_batch.hip"]
EXE_SOURCES = ["main.cpp", "putil.cpp", "boinc_stub.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_lib() -> str:
    objs = []
    for src in LIB_SOURCES:
        obj = os.path.join(PKG, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        _run([hipcc(), *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_exe() -> str:
    srcs = [os.path.join(CSRC, s) for s in EXE_SOURCES]
    tmp = EXE + ".tmp"
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", *srcs, "-o", tmp,
          f"-L{PKG}", "-lclwilson", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, EXE)
    return EXE


def build_hostcheck() -> str:
    tmp = HOSTCHECK + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas",
          f"-I{CSRC}", os.path.join(CSRC, "clw_hostcheck.cpp"), "-o", tmp])
    os.replace(tmp, HOSTCHECK)
    return HOSTCHECK


def build_all() -> None:
    build_hostcheck()
    build_lib()
    build_exe()


if __name__ == "__main__":
    build_all()
    print(LIB)
