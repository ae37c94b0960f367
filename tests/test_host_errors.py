"""Device failures reach the C caller as sticky GL errors (reference
gl_set_error, src/gl_api.c:30-35: the first error sticks until glGetError).

The CPU oracle's seam can be told to fail (oracle_inject_failure, test
infrastructure): the host GL layer -- the same sources libmtgl.so is built
from -- must map MTGL_ERR_OUT_OF_MEMORY to GL_OUT_OF_MEMORY and every other
device failure (a runtime error, a refused call) to GL_INVALID_OPERATION,
keep the first error, and keep working after it.  A batch refuses points,
lines and glDrawPixels while recording; that refusal is a
GL_INVALID_OPERATION too, on the oracle here and on the device
(tests/test_gpu_batch.py)."""
import ctypes as C

import numpy as np
import pytest

from mytinygl_amd.gl import ENUMS as E

OOM, DEVICE, UNSUPPORTED = 1, 3, 4


@pytest.fixture()
def ctx(oracle):
    oracle.lib.oracle_inject_failure.argtypes = [C.c_int, C.c_int]
    oracle.lib.oracle_inject_failure.restype = None
    c = oracle.create_context(16, 12)
    yield oracle
    oracle.lib.oracle_inject_failure(0, 0)
    oracle.destroy_context(c)
    oracle.release()


def _tri(g):
    g.glBegin(E["GL_TRIANGLES"])
    g.glVertex3f(-0.5, -0.5, 0.0)
    g.glVertex3f(0.5, -0.5, 0.0)
    g.glVertex3f(0.0, 0.5, 0.0)
    g.glEnd()


def _buffer(g):
    ids = np.zeros(1, np.uint32)
    g.glGenBuffers(1, ids)
    return int(ids[0])


def _texture(g):
    ids = np.zeros(1, np.uint32)
    g.glGenTextures(1, ids)
    return int(ids[0])


CALLS = {
    "glClear": lambda g: g.glClear(E["GL_COLOR_BUFFER_BIT"]),
    "glEnd": _tri,
    "glDrawArrays": lambda g: (g.glEnableClientState(E["GL_VERTEX_ARRAY"]),
                               g.glVertexPointer(3, E["GL_FLOAT"], 0, g.keep(np.zeros(9, np.float32))),
                               g.glDrawArrays(E["GL_TRIANGLES"], 0, 3),
                               g.glDisableClientState(E["GL_VERTEX_ARRAY"])),
    "glTexImage2D": lambda g: (g.glBindTexture(E["GL_TEXTURE_2D"], _texture(g)),
                               g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], 2, 2, 0, E["GL_RGBA"],
                                              E["GL_UNSIGNED_BYTE"], g.keep(np.zeros(16, np.uint8)))),
    "glBufferData": lambda g: (g.glBindBuffer(E["GL_ARRAY_BUFFER"], _buffer(g)),
                               g.glBufferData(E["GL_ARRAY_BUFFER"], 64, g.keep(np.zeros(64, np.uint8)),
                                              E["GL_STATIC_DRAW"])),
    "glReadPixels": lambda g: g.glReadPixels(0, 0, 4, 4, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"],
                                             g.keep(np.zeros(64, np.uint8))),
    "glDrawPixels": lambda g: (g.glRasterPos2i(0, 0),
                               g.glDrawPixels(2, 2, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"],
                                              g.keep(np.zeros(16, np.uint8)))),
    "glFinish": lambda g: g.glFinish(),
}


@pytest.mark.parametrize("call", sorted(CALLS))
@pytest.mark.parametrize("rc,gl_err", [(OOM, "GL_OUT_OF_MEMORY"), (DEVICE, "GL_INVALID_OPERATION"),
                                       (UNSUPPORTED, "GL_INVALID_OPERATION")])
def test_device_failure_sets_sticky_error(ctx, call, rc, gl_err):
    g = ctx
    assert g.glGetError() == 0
    g.lib.oracle_inject_failure(rc, 1)
    CALLS[call](g)
    g.lib.oracle_inject_failure(0, 0)
    assert g.glGetError() == E[gl_err]
    assert g.glGetError() == 0          # read once, then cleared
    CALLS[call](g)                      # the context keeps working
    assert g.glGetError() == 0


def test_first_error_sticks(ctx):
    g = ctx
    g.lib.oracle_inject_failure(OOM, 1)
    g.glClear(E["GL_COLOR_BUFFER_BIT"])
    g.lib.oracle_inject_failure(DEVICE, 1)
    _tri(g)
    g.glEnable(0x1234)                  # an invalid enum after both
    assert g.glGetError() == E["GL_OUT_OF_MEMORY"]
    assert g.glGetError() == 0


def test_failed_buffer_upload_leaves_no_buffer_storage(ctx):
    """glBufferData whose device copy fails reports GL_OUT_OF_MEMORY; the
    next upload succeeds and draws read it."""
    g = ctx
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], _buffer(g))
    g.lib.oracle_inject_failure(OOM, 1)
    g.glBufferData(E["GL_ARRAY_BUFFER"], 64, g.keep(np.zeros(64, np.uint8)), E["GL_STATIC_DRAW"])
    assert g.glGetError() == E["GL_OUT_OF_MEMORY"]
    verts = g.keep(np.array([-1, -1, 0, 1, -1, 0, 0, 1, 0], np.float32))
    g.glBufferData(E["GL_ARRAY_BUFFER"], verts.nbytes, verts, E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 0, None)
    g.glClear(E["GL_COLOR_BUFFER_BIT"])
    g.glDrawArrays(E["GL_TRIANGLES"], 0, 3)
    assert g.glGetError() == 0


def test_batch_refusals_are_errors(oracle):
    """While a batch records, points, lines, wireframe triangles and
    glDrawPixels are refused with GL_INVALID_OPERATION; filled triangles and
    clears are recorded without error."""
    g = oracle
    c = g.create_context(16, 12)
    b = g.batch()
    try:
        b.begin()
        g.glClear(E["GL_COLOR_BUFFER_BIT"])
        _tri(g)
        assert g.glGetError() == 0
        g.glBegin(E["GL_LINES"])
        g.glVertex3f(-0.5, 0.0, 0.0)
        g.glVertex3f(0.5, 0.0, 0.0)
        g.glEnd()
        assert g.glGetError() == E["GL_INVALID_OPERATION"]
        g.glPolygonMode(E["GL_FRONT_AND_BACK"], E["GL_LINE"])
        _tri(g)
        assert g.glGetError() == E["GL_INVALID_OPERATION"]
        g.glPolygonMode(E["GL_FRONT_AND_BACK"], E["GL_FILL"])
        g.glRasterPos2i(0, 0)
        g.glDrawPixels(2, 2, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], g.keep(np.zeros(16, np.uint8)))
        assert g.glGetError() == E["GL_INVALID_OPERATION"]
        b.end(autotune=False)
    finally:
        b.destroy()
        g.destroy_context(c)
        g.release()
