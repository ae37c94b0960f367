"""Build recipe of libmtgl.so (host GL layer + HIP pipeline for gfx950).

Outputs, in-tree so they travel with the repository snapshot:
  mytinygl_amd/libmtgl.so        the drop-in GL 1.5 library (C ABI)
  mytinygl_amd/_build/*.o        objects
The parity checkers are built by oracle/Makefile (see build_checkers()).

Float discipline: every translation unit is compiled without FMA contraction
and without fast-math, so fp32 expressions round exactly like the reference's
strict build.  Packed-fp32 instructions are disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmtgl.so")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("MTGL_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["gl_api.c", "gl_arrays.c", "gl_lists.c"]
HIP_SOURCES = ["mtgl_dev.hip"]

HOST_FLAGS = ["-std=c99", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fvisibility=hidden",
              "-Wall", "-Wno-unused-function", "-D_GNU_SOURCE"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... ({r.returncode})")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err and verbose:
        print(err)


def build_library(verbose: bool = False) -> str:
    """Compile every source (always; this is the build check) and link libmtgl.so."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    inc = ["-I" + INCLUDE, "-I" + CSRC]
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([hipcc, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lm", "-Wl,-Bsymbolic"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_checkers(verbose: bool = False) -> None:
    """oracle/_build/liboracle_gl.so always; oracle/_ref/libmtgl_ref.so where the
    reference sources are present."""
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-B", "all"]
    _run(cmd, verbose)


if __name__ == "__main__":
    build_library(verbose=True)
    build_checkers(verbose=True)
    print("built", LIB)
