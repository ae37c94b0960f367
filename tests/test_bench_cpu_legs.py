"""bench.py's side legs end to end on the CPU: the device library replaced by
the oracle (which exports the same entry points), the CPU baselines on the
reference renderer where oracle/_ref holds it.  Tiny sizes; this checks the
plumbing of the legs (every call they make exists on both libraries and the
records they print are well formed), not their speed."""
import os

import pytest

import bench
from mytinygl_amd import gl as GL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.environ.get("MTGL_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")


@pytest.fixture
def oracle_as_device(monkeypatch):
    if not os.path.exists(ORACLE):
        pytest.skip("oracle not built")
    real = GL.GLLibrary

    def lib(path, kind="mtgl"):
        if path.endswith(os.path.join("mytinygl_amd", "libmtgl.so")):
            return real(ORACLE, "oracle")
        return real(path, kind)
    monkeypatch.setattr(bench, "GLLibrary", lib)
    monkeypatch.setattr(bench, "W", 64)
    monkeypatch.setattr(bench, "H", 48)


def test_lines_leg(oracle_as_device):
    out = bench.lines_bench(1, 1, n_lines=200, n_points=200)
    assert out["fragments_per_frame"] > 0
    cb = out["cpu_baseline"]
    assert cb["kind"] in ("reference", "port") and cb["cpu"] and cb["ms_per_frame"] > 0


def test_lists_leg(oracle_as_device):
    out = bench.lists_bench(1, 1, copies=2)
    assert out["fragments_per_frame"] > 0
    assert out["upload_bytes_per_frame_list"] == 0
    assert out["cpu_baseline"]["kind"] in ("reference", "port")


def test_pixels_leg(oracle_as_device):
    out = bench.pixels_bench(1, 1)
    assert out["calls"] == 1
    assert out["cpu_baseline"]["kind"] in ("reference", "port")


def test_cpu_library_prefers_the_reference():
    lib, loader, kind = bench.cpu_library()
    ref = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
    if os.path.exists(ref):
        assert (lib, loader, kind) == (ref, "reference", "reference")
    else:
        assert kind in ("port", None)
