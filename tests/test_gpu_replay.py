"""Draws that overflow the device's arrays run again, in order.

The device queues draws without waiting for them.  It sizes the pair
arrays of a launch from the pairs per triangle of earlier launches and the
clip pool from earlier clipped triangles; a launch that needs more poisons
the context's stream (its raster stage and every later framebuffer write
are skipped), and the next call that waits for the stream grows the arrays
and runs the logged clears and draws again from that launch on.  These
frames make the first launches of a fresh context overflow both ways,
between other clears and draws, and must equal the oracle's frames."""
import numpy as np
import pytest

from conftest import color_histogram, plane_diffs
from mytinygl_amd import scenes as S
from mytinygl_amd.gl import ENUMS as E

pytestmark = pytest.mark.gpu


def _big_triangles(g, rng, n, W, H):
    """n triangles of about a third of the screen each: tens of bins per triangle."""
    g.glBegin(E["GL_TRIANGLES"])
    for _ in range(n):
        c = rng.uniform(0, 1, 2)
        for k in range(3):
            g.glColor4f(*rng.random(4).astype(np.float32))
            p = c + rng.uniform(-0.45, 0.45, 2)
            g.glVertex3f(float(p[0] * W), float(p[1] * H), float(-rng.uniform(0.2, 50.0)))
    g.glEnd()


def _near_plane_fan(g, rng, n):
    """n triangles crossing the near plane (each clipped into up to 2 fan triangles)."""
    g.glBegin(E["GL_TRIANGLES"])
    for _ in range(n):
        for k in range(3):
            g.glColor4f(*rng.random(4).astype(np.float32))
            x, y = rng.uniform(-1.5, 1.5, 2)
            z = rng.uniform(-4.0, 1.5)          # some in front of the eye, some behind it
            g.glVertex3f(float(x), float(y), float(z))
    g.glEnd()


def scene_overflows(g, W, H, seed):
    rng = np.random.default_rng(seed)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_BLEND"])
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
    _big_triangles(g, rng, 40, W, H)             # overflows the initial pair arrays
    g.glDisable(E["GL_BLEND"])
    g.glEnable(E["GL_SCISSOR_TEST"])
    g.glScissor(W // 4, H // 4, W // 2, H // 3)
    g.glClearColor(0.3, 0.2, 0.1, 1.0)
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
    g.glDisable(E["GL_SCISSOR_TEST"])
    _big_triangles(g, rng, 200, W, H)            # overflows again, after the first grew them
    S._perspective(g, W, H)
    g.glDepthFunc(E["GL_LEQUAL"])
    _near_plane_fan(g, rng, 12000)               # more clipped fan triangles than the clip pool holds
    g.glRasterPos2i(0, 0)
    g.glDrawPixels(4, 4, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], g.keep(np.full(64, 200, np.uint8)))
    _big_triangles(g, rng, 10, W, H)


@pytest.mark.parametrize("seed", [1, 2])
def test_overflowed_launches_run_again_in_order(mtgl, oracle, seed):
    W, H = 320, 200
    (a, ea), (b, eb) = S.render(oracle, scene_overflows, W, H, seed), S.render(mtgl, scene_overflows, W, H, seed)
    assert ea == eb
    d = plane_diffs(a, b)
    hist = color_histogram(a[0], b[0])
    print(f"overflows/{seed}: diffs {d} histogram {dict((i, int(v)) for i, v in enumerate(hist) if v)}", flush=True)
    assert d["depth"] == 0 and d["stencil"] == 0
    assert hist[2:].sum() == 0 and hist[1] <= max(1, a[0].size // 10000)
    assert (a[0] != a[0][0, 0]).sum() > W * H // 4   # the frame is not trivial


def test_statistics_count_each_launch_once(mtgl, oracle):
    """Fragment and pair counts after replays equal a clean run's."""
    W, H = 320, 200
    out = []
    for g in (oracle, mtgl):
        ctx = g.create_context(W, H)
        try:
            g.stats_reset(ctx)
            scene_overflows(g, W, H, 3)
            st = g.stats(ctx)
            out.append((st.fragments, st.triangles_raster))
        finally:
            g.destroy_context(ctx)
            g.release()
    print("oracle", out[0], "device", out[1])
    assert out[0] == out[1]
