"""Build libwcgpu.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with -ffp-contract=off (no fused
multiply-add, so the FP64 kernels round exactly like the sequential form)
and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwcgpu.so")
ARCH = os.environ.get("WC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["wc_scan.hip", "wc_pcg.hip", "wc_sweep.hip", "wc_amg.hip", "wc_particles.hip",
           "wc_grid.hip", "wc_adjoint.hip", "wc_mix.hip", "wc_capi.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_host_example(verbose)
    return LIB


HOST_EXE = os.path.join(HERE, "host", "dam_break")


def build_host_example(verbose: bool = False) -> str:
    """The C++ host mirror's example program, linked against libwcgpu.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "host"),
           os.path.join(HERE, "host", "dam_break.cpp"), "-o", HOST_EXE,
           "-L", HERE, "-lwcgpu", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return HOST_EXE


def asm(src: str) -> str:
    """Device assembly of one source (for inspecting the ISA of our kernels)."""
    out = os.path.join(OBJ, os.path.splitext(src)[0] + f"-{ARCH}.s")
    os.makedirs(OBJ, exist_ok=True)
    _run([HIPCC, *COMMON, "-x", "hip", "--cuda-device-only", "-S",
          os.path.join(CSRC, src), "-o", out])
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
