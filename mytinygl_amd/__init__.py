"""MyTinyGL's per-draw pipeline on AMD Instinct MI355X (gfx950).

The product is ``libmtgl.so`` (built in this directory by ``build.py``): the
GL 1.5 C ABI of MyTinyGL (reference include/GL/gl.h:530-669 and the context
calls of src/mytinygl.h:227-233), implemented by host C code that keeps the
reference's state machine and error semantics, and a HIP backend that runs
every draw on the GPU: vertex transform and lighting, clipping and triangle
setup, tile binning, coverage, the per-fragment pipeline and texturing.

``load()`` binds the built library through ctypes.  There is no CPU
fallback: without the library, or without a GPU, context creation fails.
"""
from __future__ import annotations

import os
import sys


def _hip_runtime_order() -> None:
    """Import torch (when it is installed) before anything loads libmtgl.so.

    torch ships its own HIP runtime under the same soname as the system one
    (libamdhip64.so.7).  The first one loaded serves the whole process: with
    torch's loaded first both torch and libmtgl.so find the GPU, while with
    libmtgl.so's loaded first torch's kernels find none.  Importing the
    package therefore imports torch first; MTGL_NO_TORCH=1 skips it (for
    processes that never use torch)."""
    if "torch" in sys.modules or os.environ.get("MTGL_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


_hip_runtime_order()

from .gl import ENUMS, GLLibrary  # noqa: E402

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmtgl.so")

__all__ = ["ENUMS", "GLLibrary", "LIB_PATH", "load"]


def load() -> GLLibrary:
    """Bind libmtgl.so; raises FileNotFoundError when it has not been built."""
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                                f"python -c 'from mytinygl_amd.build import build_library; build_library()'")
    return GLLibrary(LIB_PATH, "mtgl")
