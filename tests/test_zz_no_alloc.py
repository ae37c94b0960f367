"""mtgl_dev_draw_pixels_static (include/mtgl_dev.h): glDrawPixels through a
fixed device buffer, with no device or pinned allocation, against the
allocating entry point mtgl_dev_draw_pixels through the C ABI.

Two contexts receive the same frame; then each receives the same pixels,
one through each entry point, twice in a row (the second image smaller,
another format and state), and the planes must be identical after each
call.  A source image larger than the buffer is refused and draws nothing."""
import ctypes as C

import numpy as np
import pytest

from conftest import MTGL_LIB
from mytinygl_amd import scenes as S
from mytinygl_amd.gl import ENUMS as E, GLLibrary

pytestmark = pytest.mark.gpu

F_DEPTH_TEST, F_BLEND, F_ALPHA_TEST = 1 << 1, 1 << 3, 1 << 9
STATIC_MAX = 32 << 20


class MtglPixels(C.Structure):
    """mtgl_pixels of include/mtgl_dev.h."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("format", C.c_uint32), ("flags", C.c_uint32), ("alpha_func", C.c_uint32),
                ("alpha_ref", C.c_float), ("depth_func", C.c_uint32), ("depth_mask", C.c_uint32),
                ("blend_src", C.c_uint32), ("blend_dst", C.c_uint32), ("pixels", C.c_void_p)]


def _frame(g, W, H):
    g.glClearColor(0.2, 0.1, 0.3, 1.0)
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
    S.FEATURE_SCENES["blend"](g, W, H, 3)


def _desc(img, x, y, w, h, fmt, flags, alpha_ref=0.3):
    return MtglPixels(x, y, w, h, E[fmt], flags, E["GL_GREATER"], alpha_ref, E["GL_LEQUAL"], 1,
                      E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"], img.ctypes.data)


def test_static_draw_pixels_matches_allocating_entry_point():
    W, H = 333, 217
    g = GLLibrary(MTGL_LIB, "mtgl")
    for name in ("mtgl_dev_draw_pixels", "mtgl_dev_draw_pixels_static"):
        fn = getattr(g.lib, name)
        fn.argtypes = [C.c_void_p, C.POINTER(MtglPixels)]
        fn.restype = C.c_int
    rng = np.random.default_rng(17)
    ctxs = [g.create_context(W, H) for _ in range(2)]
    try:
        for c in ctxs:
            g.gl_make_current(c)
            _frame(g, W, H)
        calls = [
            # a large RGBA image hanging off the left and top, depth test, blending, alpha test
            (rng.integers(0, 256, (300, 320, 4), dtype=np.uint8), -20, 10, "GL_RGBA",
             F_DEPTH_TEST | F_BLEND | F_ALPHA_TEST),
            # then a smaller RGB image inside, blending only
            (rng.integers(0, 256, (37, 53, 3), dtype=np.uint8), 100, 60, "GL_RGB", F_BLEND),
        ]
        for k, (img, x, y, fmt, flags) in enumerate(calls):
            h, w = img.shape[:2]
            img = np.ascontiguousarray(img)
            d0 = _desc(img, x, y, w, h, fmt, flags)
            d1 = _desc(img, x, y, w, h, fmt, flags)
            print(f"call {k}: {fmt} {w}x{h} at ({x}, {y}) flags {flags:#x}", flush=True)
            assert g.lib.mtgl_dev_draw_pixels(g.device(ctxs[0]), C.byref(d0)) == 0
            assert g.lib.mtgl_dev_draw_pixels_static(g.device(ctxs[1]), C.byref(d1)) == 0
            a = g.read_framebuffer(ctxs[0], W, H)
            b = g.read_framebuffer(ctxs[1], W, H)
            for plane, p, q in zip(("color", "depth", "stencil"), a, b):
                assert np.array_equal(p.view(np.uint8), q.view(np.uint8)), f"call {k}: {plane} differs"
            assert (a[0] != b"\x00").any()
        # an image over the buffer's size is refused and changes nothing
        big_h = STATIC_MAX // (4 * 2048) + 1
        big = np.zeros((big_h, 2048, 4), np.uint8)
        before = g.read_framebuffer(ctxs[1], W, H)
        d = _desc(big, 0, 0, 2048, big_h, "GL_RGBA", 0)
        assert g.lib.mtgl_dev_draw_pixels_static(g.device(ctxs[1]), C.byref(d)) == 2   # MTGL_ERR_INVALID
        after = g.read_framebuffer(ctxs[1], W, H)
        assert all(np.array_equal(p, q) for p, q in zip(before, after))
    finally:
        for c in ctxs:
            g.destroy_context(c)
        g.release()
