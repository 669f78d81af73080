"""Build recipe of the channel-flow HIP library and the Fortran driver.

``build_library()`` compiles csrc/channelflow.hip for gfx950 into
lib/libchannelflow.so (the C-ABI of include/channelflow.h);
``build_driver()`` compiles the Fortran driver fortran/les_gpu.F90 against it
when a Fortran compiler is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libchannelflow.so")
SRC = os.path.join(HERE, "csrc", "channelflow.hip")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 (v_pk_*_f32) instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def _flang():
    for c in ("/opt/rocm/lib/llvm/bin/flang", "/opt/rocm/llvm/bin/flang", shutil.which("flang") or ""):
        if c and os.path.exists(c):
            return c
    return None


def build_driver(verbose: bool = False):
    """Fortran driver les_gpu (ISO_C_BINDING over the C-ABI).  Returns its
    path, or None when no Fortran compiler exists."""
    fc = _flang()
    src = os.path.join(HERE, "fortran", "les_gpu.F90")
    if fc is None or not os.path.exists(src):
        return None
    out = os.path.join(HERE, "fortran", "les_gpu")
    moddir = os.path.join(HERE, "fortran", "_mod")
    os.makedirs(moddir, exist_ok=True)
    libdir = os.path.dirname(LIB)
    cmd = [fc, "-O2", "-J", moddir, "-o", out, src, f"-L{libdir}", "-lchannelflow",
           f"-Wl,-rpath,{libdir}", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stderr[-4000:])
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build_library(verbose=True)
    build_driver(verbose=True)
