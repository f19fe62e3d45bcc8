"""Build libps3d.so for gfx950 with hipcc, in-tree (ps3d_amd/libps3d.so).

    python -m ps3d_amd.build
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
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libps3d.so")
SOURCES = ["stencil.hip", "fft.hip", "tridiag.hip", "api.hip", "ensemble.hip", "balance.hip", "monitor.hip", "tune.hip",
           "transforms.hip", "probe.hip"]
ARCH = os.environ.get("PS3D_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are kept out of the generated code
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include")]
# extra compile flags, e.g. PS3D_HIPCC_FLAGS="-DPS3D_FRIC_TY=16" (tile variants)
FLAGS += os.environ.get("PS3D_HIPCC_FLAGS", "").split()


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        for s in SOURCES:
            o = os.path.join(OBJ, s.replace(".hip", ".o"))
            jobs.append(ex.submit(_run, [HIPCC, *FLAGS, "-c", os.path.join(CSRC, s), "-o", o]))
        for j in jobs:
            out = j.result()
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *[os.path.join(OBJ, s.replace(".hip", ".o")) for s in SOURCES]])
    os.replace(tmp, LIB)
    return LIB


def build_fortran_interfaces() -> str | None:
    """Compile the ISO_C_BINDING interface module as a check (needs amdflang)."""
    fc = shutil.which("amdflang")
    if not fc:
        return None
    os.makedirs(OBJ, exist_ok=True)
    src = os.path.join(HERE, "fortran", "ps3d_capi.f90")
    obj = os.path.join(OBJ, "ps3d_capi.o")
    _run([fc, "-O2", "-fPIC", "-c", src, "-o", obj, "-module-dir", OBJ])
    return obj


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
