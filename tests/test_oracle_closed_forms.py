"""Checks of the CPU oracle that do not use the reference: each compares the
oracle with a closed form or a brute-force computation written from the
mathematics of the operation, chosen so that a plausible slip (a dropped
term, a swapped operand, an off-by-one in the fill rule or texel address)
fails.

- coverage: integer screen triangles against an int64 brute-force inclusive
  edge test (src/raster.c:299-302, :532-540), bounding box clamped to the
  framebuffer, |area| < 0.5 dropped;
- texel addressing and byte round trip: a 1:1 GL_NEAREST / GL_REPLACE quad
  reproduces the texture bytes exactly (pack(unpack(b)) == b for all bytes);
- level 1: a 2:1 minified quad under GL_NEAREST_MIPMAP_NEAREST shows the
  2x2 box filter (((c00 + c10) + c01) + c11) * 0.25 truncated to bytes;
- stencil: every function and operation against a truth table;
- depth: every comparison function against a truth table at exact depths.
"""
import itertools

import numpy as np
import pytest

from mytinygl_amd.gl import ENUMS as E

FUNCS = ["GL_NEVER", "GL_LESS", "GL_EQUAL", "GL_LEQUAL", "GL_GREATER", "GL_NOTEQUAL", "GL_GEQUAL", "GL_ALWAYS"]
OPS = ["GL_KEEP", "GL_ZERO", "GL_REPLACE", "GL_INCR", "GL_DECR", "GL_INVERT"]


def _cmp(func, a, b):
    return {"GL_NEVER": False, "GL_LESS": a < b, "GL_EQUAL": a == b, "GL_LEQUAL": a <= b, "GL_GREATER": a > b,
            "GL_NOTEQUAL": a != b, "GL_GEQUAL": a >= b, "GL_ALWAYS": True}[func]


def _sop(op, s, ref):
    return {"GL_KEEP": s, "GL_ZERO": 0, "GL_REPLACE": ref & 0xFF, "GL_INCR": min(s + 1, 255),
            "GL_DECR": max(s - 1, 0), "GL_INVERT": (~s) & 0xFF}[op]


def _ortho_ctx(g, W, H, near=-1.0, far=1.0):
    ctx = g.create_context(W, H)
    g.glViewport(0, 0, W, H)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), near, far)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    return ctx


def _done(g, ctx):
    g.destroy_context(ctx)
    g.release()


def _screen_xy(xo, yo, H):
    # ortho(0, W, 0, H) with the sample at integer pixel positions: an object
    # point (k + 0.25, m + 0.25) lands on screen (k, H - m - 1) (x truncation,
    # y flipped)
    return int(np.floor(xo)), H - int(np.floor(yo)) - 1


def _brute_coverage(tris, W, H):
    cov = np.zeros((H, W), np.int64)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.int64)
    for (x0, y0), (x1, y1), (x2, y2) in tris:
        area = (x2 - x0) * (y1 - y0) - (y2 - y0) * (x1 - x0)
        if abs(area) < 0.5:
            continue
        w0 = (xs - x1) * (y2 - y1) - (ys - y1) * (x2 - x1)
        w1 = (xs - x2) * (y0 - y2) - (ys - y2) * (x0 - x2)
        w2 = (xs - x0) * (y1 - y0) - (ys - y0) * (x1 - x0)
        inside = ((area > 0) & (w0 >= 0) & (w1 >= 0) & (w2 >= 0)) | ((area < 0) & (w0 <= 0) & (w1 <= 0) & (w2 <= 0))
        bx0, bx1 = max(min(x0, x1, x2), 0), min(max(x0, x1, x2), W - 1)
        by0, by1 = max(min(y0, y1, y2), 0), min(max(y0, y1, y2), H - 1)
        box = (xs >= bx0) & (xs <= bx1) & (ys >= by0) & (ys <= by1)
        cov += (inside & box)
    return cov


@pytest.mark.parametrize("seed", range(6))
def test_coverage_brute_force(oracle, seed):
    W, H = 48, 40
    rng = np.random.default_rng(1000 + seed)
    ctx = _ortho_ctx(oracle, W, H)
    try:
        oracle.glClearStencil(0)
        oracle.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
        oracle.glEnable(E["GL_STENCIL_TEST"])
        oracle.glStencilFunc(E["GL_ALWAYS"], 0, 0xFF)
        oracle.glStencilOp(E["GL_INCR"], E["GL_INCR"], E["GL_INCR"])
        tris = []
        oracle.glBegin(E["GL_TRIANGLES"])
        for t in range(40):
            pts = rng.integers(0, [W, H], size=(3, 2))
            if t % 7 == 0:       # collinear (area 0) and thin triangles
                pts[2] = pts[0] + (pts[1] - pts[0]) // 2
            for px, py in pts:
                xo, yo = px + 0.25, py + 0.25
                oracle.glVertex2f(float(xo), float(yo))
            tris.append([_screen_xy(px + 0.25, py + 0.25, H) for px, py in pts])
        oracle.glEnd()
        _, _, st = oracle.read_framebuffer(ctx, W, H)
    finally:
        _done(oracle, ctx)
    want = _brute_coverage(tris, W, H)
    assert want.max() < 255
    np.testing.assert_array_equal(st.astype(np.int64), want)


def _textured_quad(g, W, H, u0, u1, v0, v1):
    g.glBegin(E["GL_QUADS"])
    g.glTexCoord2f(u0, v1); g.glVertex2f(0.0, 0.0)
    g.glTexCoord2f(u1, v1); g.glVertex2f(float(W), 0.0)
    g.glTexCoord2f(u1, v0); g.glVertex2f(float(W), float(H))
    g.glTexCoord2f(u0, v0); g.glVertex2f(0.0, float(H))
    g.glEnd()


def _upload(g, tex, min_filter):
    ids = np.zeros(1, np.uint32)
    g.glGenTextures(1, ids)
    g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[0]))
    h, w, _ = tex.shape
    g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], w, h, 0, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"],
                   g.keep(np.ascontiguousarray(tex)))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E[min_filter])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_NEAREST"])
    g.glTexEnvi(E["GL_TEXTURE_ENV"], E["GL_TEXTURE_ENV_MODE"], E["GL_REPLACE"])
    g.glEnable(E["GL_TEXTURE_2D"])


def test_texel_exact_nearest(oracle):
    # screen row y samples v = (y + 0.5) / H: texel row y; column likewise
    W, H = 32, 24
    tex = np.random.default_rng(7).integers(0, 256, (H, W, 4), dtype=np.uint8)
    ctx = _ortho_ctx(oracle, W, H)
    try:
        _upload(oracle, tex, "GL_NEAREST")
        _textured_quad(oracle, W, H, 0.5 / W, (W + 0.5) / W, 0.5 / H, (H + 0.5) / H)
        c, _, _ = oracle.read_framebuffer(ctx, W, H)
    finally:
        _done(oracle, ctx)
    np.testing.assert_array_equal(c.view(np.uint8).reshape(H, W, 4), tex)


def test_level1_box_filter(oracle):
    # a 2W x 2H texture on a W x H quad: texels per pixel 4, lod = 1 >= 0.5,
    # so GL_NEAREST_MIPMAP_NEAREST reads level 1 at its texel (x, y)
    W, H = 16, 12
    tex = np.random.default_rng(11).integers(0, 256, (2 * H, 2 * W, 4), dtype=np.uint8)
    ctx = _ortho_ctx(oracle, W, H)
    try:
        _upload(oracle, tex, "GL_NEAREST_MIPMAP_NEAREST")
        _textured_quad(oracle, W, H, 0.5 / W, (W + 0.5) / W, 0.5 / H, (H + 0.5) / H)
        c, _, _ = oracle.read_framebuffer(ctx, W, H)
    finally:
        _done(oracle, ctx)
    f = tex.astype(np.float32) / np.float32(255.0)
    box = (((f[0::2, 0::2] + f[0::2, 1::2]) + f[1::2, 0::2]) + f[1::2, 1::2]) * np.float32(0.25)
    want = (np.clip(box, 0, 1) * np.float32(255.0)).astype(np.uint8)
    np.testing.assert_array_equal(c.view(np.uint8).reshape(H, W, 4), want)


@pytest.mark.parametrize("func", FUNCS)
def test_stencil_truth_table(oracle, func):
    W, H = 2, 2
    cases = list(itertools.product([0, 7, 128, 255], OPS, [(5, 0xFF, 0xFF), (300, 0x0F, 0x3C), (7, 0x80, 0xFF)]))
    for s0, op, (ref, mask, wmask) in cases:
        for sfail_first in (True, False):
            ctx = _ortho_ctx(oracle, W, H)
            try:
                oracle.glClearStencil(s0)
                oracle.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
                oracle.glEnable(E["GL_STENCIL_TEST"])
                oracle.glStencilFunc(E[func], ref, mask)
                fail_op, pass_op = (op, "GL_KEEP") if sfail_first else ("GL_KEEP", op)
                oracle.glStencilOp(E[fail_op], E["GL_KEEP"], E[pass_op])
                oracle.glStencilMask(wmask)
                oracle.glColor3f(1.0, 1.0, 1.0)
                oracle.glBegin(E["GL_TRIANGLES"])
                for x, y in ((0.25, 0.25), (2.25, 0.25), (0.25, 2.25)):
                    oracle.glVertex2f(x, y)
                oracle.glEnd()
                c, _, st = oracle.read_framebuffer(ctx, W, H)
            finally:
                _done(oracle, ctx)
            passed = _cmp(func, (ref & mask), (s0 & mask))
            new = _sop(fail_op if not passed else pass_op, s0, ref)
            want = (s0 & ~wmask | new & wmask) & 0xFF
            # pixel (0, 1) on screen is covered by the triangle
            assert int(st[1, 0]) == want, (func, s0, op, ref, mask, wmask, sfail_first)
            assert (int(c[1, 0]) == 0xFFFFFFFF) == passed


@pytest.mark.parametrize("func", FUNCS)
def test_depth_truth_table(oracle, func):
    # glOrtho(near=-1, far=1): z_ndc = -z, depth = (z_ndc + 1) * 0.5 exactly
    W, H = 2, 2
    for zo, depth in ((0.5, 0.25), (0.0, 0.5), (-0.5, 0.75)):
        for mask in (1, 0):
            ctx = _ortho_ctx(oracle, W, H)
            try:
                oracle.glClearDepth(0.5)
                oracle.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
                oracle.glEnable(E["GL_DEPTH_TEST"])
                oracle.glDepthFunc(E[func])
                oracle.glDepthMask(mask)
                oracle.glColor3f(1.0, 1.0, 1.0)
                oracle.glBegin(E["GL_TRIANGLES"])
                for x, y in ((0.25, 0.25), (2.25, 0.25), (0.25, 2.25)):
                    oracle.glVertex3f(x, y, zo)
                oracle.glEnd()
                c, d, _ = oracle.read_framebuffer(ctx, W, H)
            finally:
                _done(oracle, ctx)
            passed = _cmp(func, np.float32(depth), np.float32(0.5))
            assert (int(c[1, 0]) == 0xFFFFFFFF) == passed, (func, zo, mask)
            assert float(d[1, 0]) == (depth if (passed and mask) else 0.5), (func, zo, mask)
