"""GPU parity of the batched path: many contexts' clears and draws recorded
into one mtgl_batch and replayed from a captured graph.

Bar: the device batch against the oracle batch (test_batch_oracle.py pins
the oracle batch to plain per-context rendering) with the per-frame bar of
test_gpu_parity.py, and the full 512-scene workload bit-exact against the
device's own per-draw path.
"""
import time

import numpy as np
import pytest

from conftest import color_histogram, plane_diffs
from mytinygl_amd import scenes as S
from test_gpu_parity import _check_frame

pytestmark = pytest.mark.gpu

FEATURE = sorted(n for n in S.FEATURE_SCENES if n not in S.NOT_BATCHABLE)
SMALL = S.batch_specs(48, seed=S.SEED_BASE + 8, fb=(16, 256), tris=(20, 8000), tex=(16, 128))


def _batched(g, items, launches=1, autotune=False):
    r = S.BatchRun(g, items)
    try:
        b = r.record(autotune=autotune)
        before = [r.read(i) for i in range(len(items))]
        b.stats_reset()
        for _ in range(launches):
            b.launch()
        rc = b.sync()
        return before, [r.read(i) for i in range(len(items))], b.info(), b.stats(), rc
    finally:
        r.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_feature_scenes(mtgl, oracle, seed):
    items = S.feature_items(FEATURE, seed)
    _, ref, _, sref, _ = _batched(oracle, items)
    before, got, info, sgot, rc = _batched(mtgl, items)
    assert rc == 0
    print(f"feature batch: {info.contexts} contexts, {info.commands} commands, {info.rounds} rounds")
    for name, a, b, z in zip(FEATURE, ref, got, before):
        _check_frame(f"batch/{name}/{seed}", a, b)
        # building the batch left the new framebuffers as they were (zero)
        assert not z[0].any() and not z[1].any() and not z[2].any(), name
    assert sgot.fragments == sref.fragments


def test_batch_small_specs_autotuned(mtgl, oracle):
    items = S.batch_items(SMALL)
    _, ref, _, sref, _ = _batched(oracle, items, launches=2)
    _, got, info, sgot, rc = _batched(mtgl, items, launches=2, autotune=True)
    assert rc == 0
    print(f"small batch: rounds {info.rounds} scenes {info.scenes} heavy_bin {info.heavy_bin} "
          f"grid_div {info.grid_div} tune_ms {[round(info.tune_ms[i], 3) for i in range(info.tune_candidates)]}")
    assert info.rounds == 2 and info.scenes == len(SMALL) and info.launches == 2
    assert info.tune_candidates > 0
    for i, (a, b) in enumerate(zip(ref, got)):
        _check_frame(f"small/{i}", a, b)
    assert sgot.fragments == sref.fragments and sgot.frag_bytes == sref.frag_bytes


def test_batch_orders_context_work(mtgl, oracle):
    """Work issued on a context after a launch runs after it: a scissored
    clear right after the launch stays visible over the batch's frame."""
    items = S.batch_items(SMALL[:6])
    outs = []
    for g in (oracle, mtgl):
        r = S.BatchRun(g, items)
        try:
            b = r.record(autotune=False)
            b.launch()
            for ctx, it in zip(r.ctxs, items):
                g.gl_make_current(ctx)
                g.glEnable(S.E["GL_SCISSOR_TEST"])
                g.glScissor(0, 0, max(1, it[0] // 2), max(1, it[1] // 3))
                g.glClearColor(1.0, 0.0, 1.0, 1.0)
                g.glClear(S.E["GL_COLOR_BUFFER_BIT"])
                g.glDisable(S.E["GL_SCISSOR_TEST"])
            outs.append([r.read(i) for i in range(len(items))])
        finally:
            r.close()
    for i, (a, b) in enumerate(zip(*outs)):
        _check_frame(f"order/{i}", a, b)


def test_batch_retired_when_objects_change(mtgl):
    specs = [s for s in SMALL if s["textured"]][:2]
    r = S.BatchRun(mtgl, S.batch_items(specs))
    try:
        b = r.record(autotune=False)
        b.launch()
        assert b.sync() == 0
        mtgl.gl_make_current(r.ctxs[1])
        tex = np.zeros((4, 4, 4), np.uint8)
        mtgl.glTexImage2D(S.E["GL_TEXTURE_2D"], 0, S.E["GL_RGBA"], 4, 4, 0, S.E["GL_RGBA"],
                          S.E["GL_UNSIGNED_BYTE"], mtgl.keep(tex))
        with pytest.raises(RuntimeError):
            b.launch()
    finally:
        r.close()


def test_batch_full_workload_matches_per_draw(mtgl):
    """The 512-scene workload: one batched launch renders bit for bit what 512
    per-draw launches render, and counts the same fragments."""
    specs = S.batch_specs()
    items = S.batch_items(specs)
    r = S.BatchRun(mtgl, items)
    try:
        for ctx in r.ctxs:
            mtgl.stats_reset(ctx)
        t0 = time.perf_counter()
        r.frames_immediate()
        r.finish()
        t1 = time.perf_counter()
        frags = sum(mtgl.stats(ctx).fragments for ctx in r.ctxs)
        ref = [r.read(i) for i in range(len(items))]
        # scribble over every framebuffer so the batch has to redraw it all
        for ctx, sp in zip(r.ctxs, specs):
            mtgl.gl_make_current(ctx)
            mtgl.glClearColor(0.5, 0.5, 0.5, 0.5)
            mtgl.glClear(S.E["GL_COLOR_BUFFER_BIT"] | S.E["GL_DEPTH_BUFFER_BIT"])
            mtgl.glClearColor(sp["clear"][0], sp["clear"][1], sp["clear"][2], 1.0)
        b = r.record(autotune=True)
        b.stats_reset()
        t2 = time.perf_counter()
        b.launch()
        assert b.sync() == 0
        t3 = time.perf_counter()
        info, st = b.info(), b.stats()
        print(f"512 scenes: per-draw {1e3 * (t1 - t0):.1f} ms, batched {1e3 * (t3 - t2):.1f} ms, "
              f"fragments {frags}, heavy_bin {info.heavy_bin} grid_div {info.grid_div}", flush=True)
        assert st.fragments == frags
        bad = []
        for i in range(len(items)):
            d = plane_diffs(ref[i], r.read(i))
            if any(d.values()):
                bad.append((i, d))
        assert not bad, bad[:5]
    finally:
        r.close()
