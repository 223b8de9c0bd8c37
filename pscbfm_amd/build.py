"""Builds the HIP library in-tree: pscbfm_amd/lib/libpscbfm_amd.so (gfx950).

Run as ``python -m pscbfm_amd.build``.  The build is explicit hipcc calls,
one object per source, so a failing file is named in the error.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "libpscbfm_amd.so"
ARCH = os.environ.get("PSCBFM_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "bfm_kernels.hip",
    "resort.hip",
    "host_system.cpp",
    "device_slab.cpp",
    "updater.cpp",
    "batch.cpp",
    "server.cpp",
    "capi.cpp",
    "generators.cpp",
    "bfm_io.cpp",
    "analyzers_host.cpp",
    "analyzers.hip",
    "noalloc.hip",
]

# Packed FP32 instructions stay off in every HIP object (there is no float
# arithmetic on this path; the flag keeps the rule uniform).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [cc, "-x", "hip", *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(str(obj))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
