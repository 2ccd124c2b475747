"""Builds the native libraries of phg_amd in-tree.

    libphg_hip.so   HIP kernels + C ABI (include/phg_hip.h), gfx950 only
    libphg_grid.so  host grid / PHG numbering layer (C)

Run ``python phg_amd/build.py`` or call :func:`build`.  Outputs go to
``phg_amd/lib``.  Every HIP translation unit is compiled with packed-FP32
instructions disabled; the assembly unit additionally with
``-ffp-contract=off`` so its floating-point operations follow the reference
order exactly.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(ROOT, "_build")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]
# per-unit extra flags
HIP_UNITS = {
    "runtime.hip": [],
    "assemble.hip": ["-ffp-contract=off"],
    "krylov.hip": [],
    "reorder.hip": [],
    "dictcsr.hip": [],
    "block.hip": [],
    "eop.hip": [],
}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n  " + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose: bool = False) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")

    def compile_unit(name):
        src = os.path.join(CSRC, name)
        obj = os.path.join(BUILD, name.replace(".hip", ".o"))
        _run([HIPCC, *HIP_COMMON, *HIP_UNITS[name], "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(HIP_UNITS)) as ex:
        objs = list(ex.map(compile_unit, HIP_UNITS))
    tmp = os.path.join(LIB, "libphg_hip.so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libphg_hip.so"))

    tmp = os.path.join(LIB, "libphg_grid.so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall",
          os.path.join(HOST, "phg_grid.c"), os.path.join(HOST, "bisect.c"),
          "-lm", "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libphg_grid.so"))

    # PHG's OEM_SOLVER plugin and its serial PHG stand-ins, on the HIP library
    tmp = os.path.join(LIB, "libphg_oem_hip.so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra",
          "-Wno-unused-parameter", "-I" + os.path.join(ROOT, "include"),
          os.path.join(HOST, "phg_oem_hip.c"), os.path.join(HOST, "phg_mini.c"),
          "-L" + LIB, "-lphg_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libphg_oem_hip.so"))
    if verbose:
        print("built", os.listdir(LIB))


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
