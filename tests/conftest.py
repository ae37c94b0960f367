"""Shared fixtures: the HIP library under test and the two parity checkers.

libmtgl.so (mytinygl_amd/) is the product.  The checkers are test
infrastructure: oracle/_build/liboracle_gl.so (the host GL layer linked with
the CPU restatement of the pipeline) and oracle/_ref/libmtgl_ref.so (the
reference renderer compiled strict from its own sources, present where it was
built).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from mytinygl_amd.gl import GLLibrary  # noqa: E402

MTGL_LIB = os.path.join(ROOT, "mytinygl_amd", "libmtgl.so")
# MTGL_ORACLE_LIB selects another build of the oracle (the sanitizer build)
ORACLE_LIB = os.environ.get("MTGL_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP pipeline)")


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    return GLLibrary(ORACLE_LIB, "oracle")


@pytest.fixture(scope="session")
def reference():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference library not built (oracle/_ref/libmtgl_ref.so)")
    return GLLibrary(REF_LIB, "reference")


@pytest.fixture(scope="session")
def mtgl():
    if not os.path.exists(MTGL_LIB):
        pytest.fail("libmtgl.so is not built: run __graft_entry__.build()")
    return GLLibrary(MTGL_LIB, "mtgl")


def plane_diffs(a, b):
    """Per-plane count of differing pixels between two (colour, depth, stencil)
    triples; depth compares bit patterns."""
    out = {}
    for name, x, y in zip(("color", "depth", "stencil"), a, b):
        xv = x.view(np.uint8).reshape(x.shape + (-1,))
        yv = y.view(np.uint8).reshape(y.shape + (-1,))
        out[name] = int((xv != yv).any(-1).sum())
    return out


def color_histogram(a, b):
    """Histogram of the per-pixel maximum channel difference (0..255) between
    two RGBA8 colour planes."""
    ca = a.view(np.uint8).reshape(a.shape + (4,)).astype(np.int16)
    cb = b.view(np.uint8).reshape(b.shape + (4,)).astype(np.int16)
    mx = np.abs(ca - cb).max(-1)
    return np.bincount(mx.reshape(-1), minlength=256)
