"""The batch API's semantics on the oracle (CPU): recording clears and draws
of many contexts and launching them renders exactly what issuing the same
calls one context at a time renders, and building a batch renders nothing.
The device batch is held to the oracle batch in test_gpu_batch.py."""
import numpy as np
import pytest

from conftest import plane_diffs
from mytinygl_amd import scenes as S

SMALL = S.batch_specs(10, seed=S.SEED_BASE + 7, fb=(16, 80), tris=(20, 1500), tex=(16, 64))
FEATURE = sorted(n for n in S.FEATURE_SCENES if n not in S.NOT_BATCHABLE)


def _immediate(g, items):
    r = S.BatchRun(g, items)
    try:
        r.frames_immediate()
        return [r.read(i) for i in range(len(items))]
    finally:
        r.close()


def _batched(g, items, launches=1):
    r = S.BatchRun(g, items)
    try:
        b = r.record(autotune=False)
        before = [r.read(i) for i in range(len(items))]
        for _ in range(launches):
            b.launch()
        assert b.sync() == 0
        return before, [r.read(i) for i in range(len(items))], b.info()
    finally:
        r.close()


def test_batch_specs_shape():
    sp = S.batch_specs()
    assert len(sp) == 512
    assert all(64 <= s["w"] <= 1024 and s["w"] == s["h"] for s in sp)
    assert all(98 <= s["triangles"] <= 50_000 for s in sp)
    assert sum(s["textured"] for s in sp) == 256
    assert {s["tex"] for s in sp if s["textured"]} <= {64, 128, 256, 512}


def test_oracle_batch_equals_immediate(oracle):
    items = S.batch_items(SMALL)
    ref = _immediate(oracle, items)
    before, after, info = _batched(oracle, items)
    assert info.contexts == len(SMALL) and info.commands == 2 * len(SMALL)
    assert info.rounds == 2 and info.scenes == len(SMALL)
    for i in range(len(SMALL)):
        assert plane_diffs(ref[i], after[i]) == {"color": 0, "depth": 0, "stencil": 0}, i
        # recording and building render nothing: the new contexts are still zero
        assert not before[i][0].any() and not before[i][1].any()


def test_oracle_batch_feature_scenes(oracle):
    items = S.feature_items(FEATURE, seed=2)
    ref = _immediate(oracle, items)
    _, after, info = _batched(oracle, items, launches=2)
    assert info.launches == 2 and info.contexts == len(FEATURE)
    for name, a, b in zip(FEATURE, ref, after):
        assert plane_diffs(a, b) == {"color": 0, "depth": 0, "stencil": 0}, name


def test_oracle_batch_records_once(oracle):
    r = S.BatchRun(oracle, S.batch_items(SMALL[:2]))
    try:
        b = r.record(autotune=False)
        with pytest.raises(RuntimeError):
            b.begin()
    finally:
        r.close()
