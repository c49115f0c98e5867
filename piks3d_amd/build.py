"""Build recipe for the in-tree native libraries.

    python -m piks3d_amd.build          # libpiks3d.so (+ the Fortran host if flang exists)

libpiks3d.so is compiled by hipcc for gfx950 from csrc/ (one HIP translation
unit for the kernels and the C ABI, one C++ unit for the geometry
preprocessing).  Floating-point contraction is off so every fused operation
the reference does not do is absent (the FP64 path is compared bit for bit
with the CPU oracle), and packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpiks3d.so")
FORTRAN_DIR = os.path.join(HERE, "fortran")
HOST_EXE = os.path.join(HERE, "piks3d_gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PIKS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % " ".join(cmd[:3]))
    return r


def build_lib(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ("piks3d.hip", "serve.hip", "geometry.cpp"):
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran_host(verbose=False):
    """The drop-in Fortran host (piks3d_gpu), if a Fortran compiler exists."""
    flang = shutil.which("flang") or shutil.which("amdflang")
    src = os.path.join(FORTRAN_DIR, "piks3d_host.F90")
    if not flang or not os.path.exists(src):
        return None
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    _run([flang, "-O2", "-module-dir", objdir, "-o", HOST_EXE,
          os.path.join(FORTRAN_DIR, "piks_c.F90"), src,
          "-L" + HERE, "-Wl,-rpath,$ORIGIN", "-lpiks3d"], verbose)
    return HOST_EXE


def build(verbose=False):
    build_lib(verbose)
    build_fortran_host(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
