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

from .gl import ENUMS, GLLibrary

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmtgl.so")

__all__ = ["ENUMS", "GLLibrary", "LIB_PATH", "load"]


def load() -> GLLibrary:
    """Bind libmtgl.so; raises FileNotFoundError when it has not been built."""
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                                f"python -c 'from mytinygl_amd.build import build_library; build_library()'")
    return GLLibrary(LIB_PATH, "mtgl")
