"""GL API call scripts whose observable results (glGetError codes, glGet*
values, glIs* answers) are compared between the host GL layer and the
reference: error semantics of src/gl_api.c:26-35 (sticky first error,
no-context no-ops), the matrix stacks (src/gl_api.c:463-560), enum and range
validation, display-list bookkeeping (src/lists.c) and state queries
(src/gl_api.c:1375-1552).  Each script returns a list of JSON-able values.
"""
import ctypes as C

import numpy as np

from mytinygl_amd.gl import ENUMS as E


def _ints(g, pname, n=16):
    buf = np.full(n, -12345, np.int32)
    g.glGetIntegerv(E[pname], buf)
    return [int(v) for v in buf]


def _floats(g, pname, n=16):
    buf = np.full(n, -12345.0, np.float32)
    g.glGetFloatv(E[pname], buf)
    return [float(v) for v in buf]


def _err(g):
    return int(g.glGetError())


def script_errors(g):
    out = []
    g.glEnable(0x1234); out.append(_err(g))
    g.glEnable(0x1234); g.glMatrixMode(0x9999); out.append(_err(g))
    out.append(_err(g))
    g.glEnd(); out.append(_err(g))
    g.glBegin(E["GL_TRIANGLES"]); g.glBegin(E["GL_TRIANGLES"]); out.append(_err(g))
    g.glEnd(); out.append(_err(g))
    g.glBegin(0x7777); out.append(_err(g))
    g.glViewport(0, 0, -1, 10); out.append(_err(g))
    g.glMatrixMode(E["GL_MODELVIEW"])
    for i in range(40):
        g.glPushMatrix()
        e = _err(g)
        if e:
            out.append((i, e))
            break
    for i in range(45):
        g.glPopMatrix()
        e = _err(g)
        if e:
            out.append((i, e))
            break
    g.glMatrixMode(E["GL_PROJECTION"])
    for i in range(10):
        g.glPushMatrix()
        e = _err(g)
        if e:
            out.append(("proj", i, e))
            break
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glDepthFunc(0x1); out.append(_err(g))
    g.glBlendFunc(E["GL_ONE"], 0x4242); out.append(_err(g))
    g.glStencilOp(E["GL_KEEP"], 0x1, E["GL_KEEP"]); out.append(_err(g))
    g.glCullFace(E["GL_FRONT_AND_BACK"]); out.append(_err(g))
    g.glShadeModel(0x1); out.append(_err(g))
    g.glPolygonMode(E["GL_FRONT"], E["GL_FILL"]); out.append(_err(g))
    g.glAlphaFunc(0x1, 0.5); out.append(_err(g))
    g.glLineWidth(-1.0); out.append(_err(g))
    g.glPointSize(0.0); out.append(_err(g))
    g.glScissor(0, 0, -5, 5); out.append(_err(g))
    g.glClear(0x1); out.append(_err(g))
    ids = np.zeros(4, np.uint32)
    g.glGenTextures(-1, ids); out.append(_err(g))
    g.glBindTexture(E["GL_TEXTURE_2D"], 0); out.append(_err(g))
    g.glTexParameteri(E["GL_TEXTURE_2D"], 0x4444, 0); out.append(_err(g))
    g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], -1, 4, 0, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], None)
    out.append(_err(g))
    g.glDrawArrays(0x7777, 0, 3); out.append(_err(g))
    g.glDrawArrays(E["GL_TRIANGLES"], 0, -3); out.append(_err(g))
    g.glBindBuffer(0x5555, 1); out.append(_err(g))
    g.glBufferData(E["GL_ARRAY_BUFFER"], -4, None, E["GL_STATIC_DRAW"]); out.append(_err(g))
    out.append(_err(g))
    return out


def script_matrices(g):
    out = []
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glTranslatef(1.0, 2.0, 3.0)
    g.glRotatef(33.0, 0.2, 0.7, 0.1)
    g.glScalef(2.0, 0.5, -1.5)
    out.append(_floats(g, "GL_MODELVIEW_MATRIX"))
    g.glPushMatrix()
    g.glRotatef(-71.0, 1.0, 0.0, 0.0)
    out.append(_floats(g, "GL_MODELVIEW_MATRIX"))
    g.glPopMatrix()
    out.append(_floats(g, "GL_MODELVIEW_MATRIX"))
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glFrustum(-0.1, 0.1, -0.075, 0.075, 0.1, 100.0)
    out.append(_floats(g, "GL_PROJECTION_MATRIX"))
    g.glOrtho(0.0, 640.0, 0.0, 480.0, -1.0, 1.0)
    out.append(_floats(g, "GL_PROJECTION_MATRIX"))
    out.append(_ints(g, "GL_MATRIX_MODE", 1))
    out.append(_err(g))
    return out


def script_state_queries(g):
    out = []
    g.glViewport(3, 4, 50, 60)
    g.glScissor(1, 2, 30, 40)
    g.glDepthFunc(E["GL_GEQUAL"])
    g.glDepthMask(0)
    g.glDepthRange(0.25, 0.75)
    g.glClearColor(0.1, 0.2, 0.3, 0.4)
    g.glClearDepth(0.5)
    g.glClearStencil(7)
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
    g.glStencilFunc(E["GL_LEQUAL"], 5, 0x0F)
    g.glStencilOp(E["GL_INCR"], E["GL_DECR"], E["GL_INVERT"])
    g.glStencilMask(0x3C)
    g.glAlphaFunc(E["GL_GREATER"], 0.3)
    g.glCullFace(E["GL_FRONT"])
    g.glFrontFace(E["GL_CW"])
    g.glShadeModel(E["GL_FLAT"])
    g.glColorMask(1, 0, 1, 0)
    g.glColor4f(0.25, 0.5, 0.75, 1.0)
    g.glNormal3f(0.0, 0.6, 0.8)
    g.glTexCoord2f(0.125, 0.875)
    g.glFogi(E["GL_FOG_MODE"], E["GL_LINEAR"])
    g.glFogf(E["GL_FOG_START"], 2.0)
    g.glFogf(E["GL_FOG_END"], 9.0)
    g.glFogf(E["GL_FOG_DENSITY"], 0.35)
    g.glHint(E["GL_PERSPECTIVE_CORRECTION_HINT"], E["GL_FASTEST"])
    for cap in ("GL_DEPTH_TEST", "GL_BLEND", "GL_CULL_FACE", "GL_FOG", "GL_LIGHTING", "GL_TEXTURE_2D",
                "GL_STENCIL_TEST", "GL_ALPHA_TEST", "GL_SCISSOR_TEST", "GL_LIGHT0", "GL_LIGHT3",
                "GL_COLOR_MATERIAL", "GL_NORMALIZE"):
        out.append((cap, int(g.glIsEnabled(E[cap]))))
        g.glEnable(E[cap])
        out.append((cap, int(g.glIsEnabled(E[cap]))))
    for pname, n in (("GL_VIEWPORT", 4), ("GL_SCISSOR_BOX", 4), ("GL_DEPTH_FUNC", 1), ("GL_BLEND_SRC", 1),
                     ("GL_BLEND_DST", 1), ("GL_STENCIL_FUNC", 1), ("GL_STENCIL_REF", 1),
                     ("GL_STENCIL_VALUE_MASK", 1), ("GL_STENCIL_FAIL", 1), ("GL_STENCIL_PASS_DEPTH_FAIL", 1),
                     ("GL_STENCIL_PASS_DEPTH_PASS", 1), ("GL_STENCIL_WRITEMASK", 1), ("GL_CULL_FACE_MODE", 1),
                     ("GL_FRONT_FACE", 1), ("GL_SHADE_MODEL", 1), ("GL_MATRIX_MODE", 1), ("GL_MAX_TEXTURE_SIZE", 1),
                     ("GL_MAX_LIGHTS", 1), ("GL_MAX_MODELVIEW_STACK_DEPTH", 1), ("GL_MAX_PROJECTION_STACK_DEPTH", 1),
                     ("GL_FOG_MODE", 1), ("GL_ALPHA_TEST_FUNC", 1), ("GL_STENCIL_CLEAR_VALUE", 1),
                     ("GL_DEPTH_WRITEMASK", 1), ("GL_COLOR_WRITEMASK", 4), ("GL_PERSPECTIVE_CORRECTION_HINT", 1)):
        if pname in E:
            out.append((pname, _ints(g, pname, n)))
            out.append((pname + "/err", _err(g)))
    for pname, n in (("GL_COLOR_CLEAR_VALUE", 4), ("GL_DEPTH_CLEAR_VALUE", 1), ("GL_DEPTH_RANGE", 2),
                     ("GL_CURRENT_COLOR", 4), ("GL_CURRENT_NORMAL", 3), ("GL_CURRENT_TEXTURE_COORDS", 4),
                     ("GL_FOG_START", 1), ("GL_FOG_END", 1), ("GL_FOG_DENSITY", 1), ("GL_FOG_COLOR", 4),
                     ("GL_ALPHA_TEST_REF", 1), ("GL_LINE_WIDTH", 1), ("GL_POINT_SIZE", 1)):
        if pname in E:
            out.append((pname, _floats(g, pname, n)))
            out.append((pname + "/err", _err(g)))
    out.append(_err(g))
    return out


def script_lights_materials(g):
    out = []
    g.glLightfv(E["GL_LIGHT1"], E["GL_POSITION"], g.fv(1.0, 2.0, 3.0, 1.0))
    g.glLightfv(E["GL_LIGHT1"], E["GL_DIFFUSE"], g.fv(0.5, 0.25, 0.125, 1.0))
    g.glLightf(E["GL_LIGHT1"], E["GL_SPOT_CUTOFF"], 30.0)
    g.glLightf(E["GL_LIGHT1"], E["GL_SPOT_CUTOFF"], 120.0)
    out.append(_err(g))
    g.glLightf(E["GL_LIGHT1"], E["GL_CONSTANT_ATTENUATION"], -1.0)
    out.append(_err(g))
    for pname, n in (("GL_POSITION", 4), ("GL_DIFFUSE", 4), ("GL_AMBIENT", 4), ("GL_SPECULAR", 4),
                     ("GL_SPOT_DIRECTION", 3), ("GL_SPOT_CUTOFF", 1), ("GL_SPOT_EXPONENT", 1)):
        buf = np.full(4, -7.0, np.float32)
        g.glGetLightfv(E["GL_LIGHT1"], E[pname], buf)
        out.append((pname, [float(v) for v in buf[:n]]))
    g.glMaterialf(E["GL_FRONT"], E["GL_SHININESS"], 200.0)
    out.append(_err(g))
    g.glMaterialfv(E["GL_FRONT_AND_BACK"], E["GL_EMISSION"], g.fv(0.1, 0.2, 0.3, 0.4))
    for face in ("GL_FRONT", "GL_BACK"):
        for pname in ("GL_EMISSION", "GL_AMBIENT", "GL_DIFFUSE", "GL_SPECULAR", "GL_SHININESS"):
            buf = np.full(4, -7.0, np.float32)
            g.glGetMaterialfv(E[face], E[pname], buf)
            out.append((face, pname, [float(v) for v in buf]))
    out.append(_err(g))
    return out


def script_lists(g):
    out = []
    base = int(g.glGenLists(3))
    out.append(base)
    out.append([int(g.glIsList(base + i)) for i in range(4)])
    g.glNewList(base, E["GL_COMPILE"])
    g.glColor3f(1.0, 0.0, 0.0)
    g.glNewList(base + 1, E["GL_COMPILE"])
    out.append(_err(g))
    g.glEndList()
    out.append([int(g.glIsList(base + i)) for i in range(4)])
    g.glEndList()
    out.append(_err(g))
    g.glCallList(base + 50)
    out.append(_err(g))
    g.glDeleteLists(base, 2)
    out.append([int(g.glIsList(base + i)) for i in range(4)])
    g.glNewList(0, E["GL_COMPILE"])
    out.append(_err(g))
    g.glNewList(base + 2, 0x1234)
    out.append(_err(g))
    out.append(int(g.glGenLists(0)))
    out.append(_err(g))
    return out


def script_textures_buffers(g):
    out = []
    ids = np.zeros(3, np.uint32)
    g.glGenTextures(3, ids)
    out.append([int(v) for v in ids])
    out.append([int(g.glIsTexture(int(v))) for v in ids])
    g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[1]))
    out.append([int(g.glIsTexture(int(v))) for v in ids])
    out.append(_ints(g, "GL_TEXTURE_BINDING_2D", 1) if "GL_TEXTURE_BINDING_2D" in E else None)
    g.glDeleteTextures(1, ids[1:2].copy())
    out.append([int(g.glIsTexture(int(v))) for v in ids])
    b = np.zeros(2, np.uint32)
    g.glGenBuffers(2, b)
    out.append([int(v) for v in b])
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(b[0]))
    out.append([int(g.glIsBuffer(int(v))) for v in b])
    data = g.keep(np.arange(16, dtype=np.float32))
    g.glBufferData(E["GL_ARRAY_BUFFER"], data.nbytes, data, E["GL_STATIC_DRAW"])
    out.append(_err(g))
    g.glBufferSubData(E["GL_ARRAY_BUFFER"], 60, 8, data)
    out.append(_err(g))
    g.glBufferSubData(E["GL_ARRAY_BUFFER"], 8, 8, data)
    out.append(_err(g))
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], 0)
    g.glBufferSubData(E["GL_ARRAY_BUFFER"], 0, 4, data)
    out.append(_err(g))
    g.glDeleteBuffers(2, b)
    out.append([int(g.glIsBuffer(int(v))) for v in b])
    out.append(_err(g))
    return out


def script_no_context(g):
    """Calls with no current context are silent no-ops (src/gl_api.c:26-27)."""
    g.gl_make_current(None)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glBegin(E["GL_TRIANGLES"])
    g.glVertex3f(0.0, 0.0, 0.0)
    g.glEnd()
    g.glClear(E["GL_COLOR_BUFFER_BIT"])
    return [int(g.glGetError()), int(g.glIsEnabled(E["GL_DEPTH_TEST"]))]


def _read(g, x, y, w, h, fmt, typ="GL_UNSIGNED_BYTE", room=None):
    """glReadPixels into a buffer pre-filled with 0xCD; returns the SHA-256 of
    the whole buffer (bytes past the window show what was written) and the
    count of bytes changed."""
    import hashlib
    n = room if room is not None else max(0, w) * max(0, h) * 4 + 16
    buf = np.full(max(n, 16), 0xCD, np.uint8)
    g.glReadPixels(x, y, w, h, E[fmt], E[typ], buf)
    return [hashlib.sha256(buf.tobytes()).hexdigest()[:24], int((buf != 0xCD).sum()), _err(g)]


def script_read_pixels(g, W=32, H=24):
    """glReadPixels (src/gl_api.c:1180-1229) on a frame of overlapping
    triangles with per-vertex alpha: RGBA and RGB, windows inside, past every
    edge, wholly outside, larger than the framebuffer, one pixel at each
    corner, zero and negative sizes, and formats and types that read nothing."""
    rng = np.random.default_rng(11)
    g.glClearColor(0.25, 0.5, 0.75, 0.125)
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), -1.0, 1.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glBegin(E["GL_TRIANGLES"])
    for _ in range(12):
        for _ in range(3):
            r, gg, b, a = (float(v) for v in rng.random(4))
            g.glColor4f(r, gg, b, a)
            g.glVertex2f(float(rng.uniform(-4, W + 4)), float(rng.uniform(-4, H + 4)))
    g.glEnd()
    out = [_err(g)]
    windows = [
        (0, 0, W, H), (3, 2, 10, 7), (-5, -3, 12, 9), (W - 6, H - 4, 11, 9), (-7, H - 3, W + 14, 6),
        (W - 2, -6, 9, H + 12), (-3, -3, W + 6, H + 6), (W, 0, 5, 4), (-5, 0, 5, 4), (0, H, 4, 3),
        (0, -3, 4, 3), (W + 10, H + 10, 3, 3), (-100, -100, 3, 2), (0, 0, 1, 1), (W - 1, 0, 1, 1),
        (0, H - 1, 1, 1), (W - 1, H - 1, 1, 1), (5, 5, W, 1), (5, 5, 1, H),
    ]
    for fmt in ("GL_RGBA", "GL_RGB"):
        for (x, y, w, h) in windows:
            out.append(_read(g, x, y, w, h, fmt))
    # nothing written: zero sizes, a negative height, other formats and types
    out.append(_read(g, 0, 0, 0, 4, "GL_RGBA", room=64))
    out.append(_read(g, 0, 0, 4, 0, "GL_RGB", room=64))
    out.append(_read(g, 0, 0, 4, -2, "GL_RGBA", room=64))
    out.append(_read(g, 0, 0, 4, 4, "GL_LUMINANCE", room=64))
    out.append(_read(g, 0, 0, 4, 4, "GL_RGBA", "GL_FLOAT", room=64))
    return out


SCRIPTS = {
    "read_pixels": script_read_pixels,
    "errors": script_errors,
    "matrices": script_matrices,
    "state_queries": script_state_queries,
    "lights_materials": script_lights_materials,
    "lists": script_lists,
    "textures_buffers": script_textures_buffers,
    "no_context": script_no_context,
}


def run_script(g, name, W=32, H=24):
    ctx = g.create_context(W, H)
    try:
        res = SCRIPTS[name](g)
    finally:
        g.gl_make_current(ctx)
        g.destroy_context(ctx)
        g.release()
    return res
