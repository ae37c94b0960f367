"""Device-resident display-list geometry (the host layer's list segments):
a replayed glBegin/glEnd block sends its vertex records to the device once
and draws from the segment afterwards, sends them again when they change
(a block that inherits the current colour), and renders exactly what
immediate mode renders.  Checked on the CPU oracle, which implements the
same segment entry points as libmtgl.so; the device is held to the oracle
by the lists scenes of the GPU parity tests."""
import numpy as np
import pytest

from conftest import plane_diffs
from mytinygl_amd.gl import ENUMS as E
from mytinygl_amd import scenes as S

REC = 52          # bytes of one vertex record: position, colour, texcoord, normal, state index


def _tris(g, rng, n, colour=True):
    for _ in range(3 * n):
        if colour:
            g.glColor4f(*rng.random(4).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 2.0).astype(np.float32))


def _compile(g, lid, seed, n, colour=True):
    rng = np.random.default_rng(seed)
    g.glNewList(lid, E["GL_COMPILE"])
    g.glBegin(E["GL_TRIANGLES"])
    _tris(g, rng, n, colour)
    g.glEnd()
    g.glEndList()


def test_replay_uploads_once(oracle):
    g = oracle
    ctx = g.create_context(64, 48)
    try:
        lid = g.glGenLists(1)
        _compile(g, lid, 1, 50)
        g.stats_reset(ctx)
        g.glCallList(lid)
        first = g.stats(ctx).upload_bytes
        assert first == 150 * REC
        for _ in range(5):
            g.glCallList(lid)
        assert g.stats(ctx).upload_bytes == first
        assert g.stats(ctx).draws == 6
    finally:
        g.destroy_context(ctx)


def test_inherited_colour_resends(oracle):
    g = oracle
    ctx = g.create_context(64, 48)
    try:
        lid = g.glGenLists(1)
        _compile(g, lid, 2, 10, colour=False)
        g.stats_reset(ctx)
        g.glColor3f(1.0, 0.0, 0.0)
        g.glCallList(lid)
        a = g.stats(ctx).upload_bytes
        g.glCallList(lid)
        assert g.stats(ctx).upload_bytes == a          # same records
        g.glColor3f(0.0, 1.0, 0.0)
        g.glCallList(lid)
        assert g.stats(ctx).upload_bytes == 2 * a      # colour changed: records sent again
    finally:
        g.destroy_context(ctx)


def test_recompile_and_immediate_equal(oracle):
    """Lists replayed under changing matrices and recompiled render what the
    same calls issued in immediate mode render."""
    def with_lists(g, W, H, seed):
        S._perspective(g, W, H)
        g.glEnable(E["GL_DEPTH_TEST"])
        lid = g.glGenLists(1)
        _compile(g, lid, seed, 8)
        for k in range(4):
            g.glRotatef(20.0, 0.0, 0.0, 1.0)
            g.glCallList(lid)
            if k == 1:
                _compile(g, lid, seed + 100, 8)

    def immediate(g, W, H, seed):
        S._perspective(g, W, H)
        g.glEnable(E["GL_DEPTH_TEST"])
        for k in range(4):
            g.glRotatef(20.0, 0.0, 0.0, 1.0)
            rng = np.random.default_rng(seed if k < 2 else seed + 100)
            g.glBegin(E["GL_TRIANGLES"])
            _tris(g, rng, 8)
            g.glEnd()

    a, ea = S.render(oracle, with_lists, 96, 72, 7)
    b, eb = S.render(oracle, immediate, 96, 72, 7)
    assert ea == eb == 0
    assert plane_diffs(a, b) == {"color": 0, "depth": 0, "stencil": 0}
    assert (a[0] != a[0][0, 0]).sum() > 100


def _state(g):
    out = []
    for pname, n in (("GL_CURRENT_COLOR", 4), ("GL_CURRENT_NORMAL", 3), ("GL_CURRENT_TEXTURE_COORDS", 4)):
        v = np.zeros(4, np.float32)
        g.glGetFloatv(E[pname], v)
        out.append(v[:n].copy())
    for face in ("GL_FRONT", "GL_BACK"):
        for pname in ("GL_AMBIENT", "GL_DIFFUSE"):
            v = np.zeros(4, np.float32)
            g.glGetMaterialfv(E[face], E[pname], v)
            out.append(v.copy())
    return np.concatenate(out)


@pytest.mark.parametrize("seed", [3, 11])
def test_repeated_replays_equal_calls(oracle, seed):
    """Replays whose calls the host skips (same inherited attributes as the
    previous replay) leave the current colour, normal, texcoord and the
    colour-material writes exactly as the calls do, and render the same
    frame; checked after every step of the repeated-list program."""
    states = {}
    frames = {}
    for inline in (False, True):
        rec = []
        frames[inline], err = S.render(
            oracle, lambda g, W, H, s: S.lists_repeat(g, W, H, s, inline=inline,
                                                      probe=lambda g: rec.append(_state(g))),
            96, 72, seed)
        assert err == 0
        states[inline] = np.stack(rec)
    assert np.array_equal(states[False].view(np.uint32), states[True].view(np.uint32))
    assert plane_diffs(frames[False], frames[True]) == {"color": 0, "depth": 0, "stencil": 0}
    assert (frames[False][0] != frames[False][0][0, 0]).sum() > 200
