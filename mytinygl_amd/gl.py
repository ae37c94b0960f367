"""ctypes binding of the GL 1.5 ABI exported by libmtgl.so.

The same class binds any library that exports the MyTinyGL C ABI
(reference include/GL/gl.h:530-669 plus the context calls of
src/mytinygl.h:227-233), so one scene script can drive the HIP library and
the test-only checkers identically.  Enumerant values are parsed from
include/GL/gl.h, the header C programs compile against.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(_HERE)
GL_HEADER = os.path.join(REPO_ROOT, "include", "GL", "gl.h")


def _parse_enums(path: str) -> dict:
    out = {}
    with open(path) as f:
        for m in re.finditer(r"#define\s+(GL_\w+)\s+(0x[0-9A-Fa-f]+|\d+)", f.read()):
            out.setdefault(m.group(1), int(m.group(2), 0))
    return out


ENUMS = _parse_enums(GL_HEADER)
globals().update(ENUMS)

# ---- argument type codes ---------------------------------------------------
_T = {
    "e": C.c_uint,     # GLenum / GLbitfield / GLuint
    "i": C.c_int,      # GLint / GLsizei
    "f": C.c_float,    # GLfloat / GLclampf
    "d": C.c_double,   # GLdouble / GLclampd
    "b": C.c_ubyte,    # GLboolean / GLubyte
    "l": C.c_long,     # GLsizeiptr / GLintptr
    "p": C.c_void_p,   # any pointer
}

# name: (return code or "", argument codes)
_SIGS = {
    "glEnable": ("", "e"), "glDisable": ("", "e"), "glClear": ("", "e"),
    "glClearColor": ("", "ffff"), "glClearDepth": ("", "d"), "glViewport": ("", "iiii"),
    "glMatrixMode": ("", "e"), "glLoadIdentity": ("", ""), "glPushMatrix": ("", ""),
    "glPopMatrix": ("", ""), "glOrtho": ("", "dddddd"), "glFrustum": ("", "dddddd"),
    "glTranslatef": ("", "fff"), "glRotatef": ("", "ffff"), "glScalef": ("", "fff"),
    "glMultMatrixf": ("", "p"), "glLoadMatrixf": ("", "p"),
    "glBegin": ("", "e"), "glEnd": ("", ""), "glVertex2f": ("", "ff"), "glVertex3f": ("", "fff"),
    "glVertex2i": ("", "ii"), "glVertex3i": ("", "iii"), "glColor3f": ("", "fff"),
    "glColor4f": ("", "ffff"), "glColor3ub": ("", "bbb"), "glColor4ub": ("", "bbbb"),
    "glTexCoord2f": ("", "ff"), "glNormal3f": ("", "fff"),
    "glGenTextures": ("", "ip"), "glDeleteTextures": ("", "ip"), "glBindTexture": ("", "ee"),
    "glTexImage2D": ("", "eiiiiieep"), "glTexParameteri": ("", "eei"),
    "glReadPixels": ("", "iiiieep"), "glDrawPixels": ("", "iieep"), "glRasterPos2i": ("", "ii"),
    "glRasterPos2f": ("", "ff"), "glRasterPos3f": ("", "fff"),
    "glFogi": ("", "ei"), "glFogf": ("", "ef"), "glFogfv": ("", "ep"),
    "glLightf": ("", "eef"), "glLighti": ("", "eei"), "glLightfv": ("", "eep"), "glLightiv": ("", "eep"),
    "glGetLightfv": ("", "eep"), "glMaterialf": ("", "eef"), "glMateriali": ("", "eei"),
    "glMaterialfv": ("", "eep"), "glMaterialiv": ("", "eep"), "glGetMaterialfv": ("", "eep"),
    "glLightModelf": ("", "ef"), "glLightModeli": ("", "ei"), "glLightModelfv": ("", "ep"),
    "glLightModeliv": ("", "ep"), "glColorMaterial": ("", "ee"), "glShadeModel": ("", "e"),
    "glGenBuffers": ("", "ip"), "glDeleteBuffers": ("", "ip"), "glBindBuffer": ("", "ee"),
    "glBufferData": ("", "elpe"), "glBufferSubData": ("", "ellp"),
    "glEnableClientState": ("", "e"), "glDisableClientState": ("", "e"),
    "glVertexPointer": ("", "ieip"), "glColorPointer": ("", "ieip"), "glTexCoordPointer": ("", "ieip"),
    "glNormalPointer": ("", "eip"), "glDrawArrays": ("", "eii"), "glDrawElements": ("", "eiep"),
    "glFlush": ("", ""), "glFinish": ("", ""), "glGetError": ("e", ""), "glBlendFunc": ("", "ee"),
    "glGetIntegerv": ("", "ep"), "glGetFloatv": ("", "ep"), "glGetDoublev": ("", "ep"),
    "glGetBooleanv": ("", "ep"), "glIsEnabled": ("b", "e"), "glGetString": ("p", "e"),
    "glCullFace": ("", "e"), "glFrontFace": ("", "e"), "glDepthFunc": ("", "e"), "glDepthMask": ("", "b"),
    "glAlphaFunc": ("", "ef"), "glHint": ("", "ee"), "glLineWidth": ("", "f"), "glPointSize": ("", "f"),
    "glPolygonMode": ("", "ee"), "glScissor": ("", "iiii"), "glDepthRange": ("", "dd"),
    "glColorMask": ("", "bbbb"), "glPixelStorei": ("", "ei"), "glStencilFunc": ("", "eie"),
    "glStencilOp": ("", "eee"), "glStencilMask": ("", "e"), "glClearStencil": ("", "i"),
    "glTexEnvi": ("", "eei"), "glTexEnvf": ("", "eef"), "glTexEnvfv": ("", "eep"),
    "glIsTexture": ("b", "e"), "glIsBuffer": ("b", "e"),
    "glGenLists": ("e", "i"), "glDeleteLists": ("", "ei"), "glNewList": ("", "ee"), "glEndList": ("", ""),
    "glCallList": ("", "e"), "glCallLists": ("", "iep"), "glListBase": ("", "e"), "glIsList": ("b", "e"),
    "gl_create_context": ("p", "ii"), "gl_destroy_context": ("", "p"), "gl_make_current": ("", "p"),
    "gl_get_current_context": ("p", ""),
}

GL_FUNCTIONS = tuple(n for n in _SIGS if n.startswith("gl") and not n.startswith("gl_"))


class MtglDevStats(C.Structure):
    """mtgl_dev_stats of include/mtgl_dev.h."""
    _fields_ = [
        ("fragments", C.c_uint64), ("frag_bytes", C.c_uint64), ("triangles_in", C.c_uint64),
        ("triangles_raster", C.c_uint64), ("bin_pairs", C.c_uint64), ("draws", C.c_uint64),
        ("raster_ms", C.c_double), ("raster_launches", C.c_uint64), ("upload_bytes", C.c_uint64),
    ]


class MtglBatchInfo(C.Structure):
    """mtgl_batch_info_t of include/mtgl_batch.h."""
    _fields_ = [
        ("contexts", C.c_uint32), ("commands", C.c_uint32), ("rounds", C.c_uint32), ("scenes", C.c_uint32),
        ("launches", C.c_uint32), ("heavy_bin", C.c_uint32), ("grid_div", C.c_uint32),
        ("tune_candidates", C.c_uint32), ("tune_ms", C.c_double * 8),
    ]


class Batch:
    """An mtgl_batch of include/mtgl_batch.h: record() returns a context manager
    under which this thread's clears and draws are recorded."""

    def __init__(self, gl: "GLLibrary"):
        self.gl = gl
        lib = gl.lib
        h = C.c_void_p()
        rc = lib.mtgl_batch_create(C.byref(h))
        if rc != 0:
            raise RuntimeError(f"{gl.kind}: mtgl_batch_create failed ({rc})")
        self.h = h.value

    def begin(self) -> None:
        self._check(self.gl.lib.mtgl_batch_begin(self.h), "begin")

    def end(self, autotune: bool = True) -> None:
        self._check(self.gl.lib.mtgl_batch_end(self.h, 1 if autotune else 0), "end")

    def launch(self) -> None:
        self._check(self.gl.lib.mtgl_batch_launch(self.h), "launch")

    def sync(self) -> int:
        return self.gl.lib.mtgl_batch_sync(self.h)

    def info(self) -> MtglBatchInfo:
        i = MtglBatchInfo()
        self._check(self.gl.lib.mtgl_batch_info(self.h, C.byref(i)), "info")
        return i

    def stats(self) -> MtglDevStats:
        s = MtglDevStats()
        self._check(self.gl.lib.mtgl_batch_stats_get(self.h, C.byref(s)), "stats")
        return s

    def stats_reset(self) -> None:
        self._check(self.gl.lib.mtgl_batch_stats_reset(self.h), "stats_reset")

    def destroy(self) -> None:
        if self.h:
            self.gl.lib.mtgl_batch_destroy(self.h)
            self.h = None

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise RuntimeError(f"{self.gl.kind}: mtgl_batch_{what} failed ({rc})")


def _ptr(a) -> Optional[int]:
    """Pointer argument: None, an int address, bytes, a ctypes object or a numpy array."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array arguments must be C-contiguous")
        return a.ctypes.data
    if isinstance(a, (bytes, bytearray)):
        return C.cast(C.c_char_p(bytes(a)), C.c_void_p).value
    return C.cast(a, C.c_void_p).value


class GLLibrary:
    """Bound GL library.  kind is "mtgl" (the HIP library), "oracle" or "reference"."""

    def __init__(self, path: str, kind: str = "mtgl"):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.kind = kind
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        self._keep = []
        for name, (ret, args) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.restype = _T[ret] if ret else None
            fn.argtypes = [_T[a] for a in args]
            setattr(self, name, self._wrap(fn, args))
        if kind == "reference":
            self._read = self.lib.mtgl_ref_read_framebuffer
        else:
            self._read = self.lib.mtgl_read_framebuffer
            self.lib.mtgl_context_device.restype = C.c_void_p
            self.lib.mtgl_context_device.argtypes = [C.c_void_p]
            self.lib.mtgl_dev_stats_get.argtypes = [C.c_void_p, C.POINTER(MtglDevStats)]
            self.lib.mtgl_dev_stats_reset.argtypes = [C.c_void_p]
            self.lib.mtgl_dev_set_timing.argtypes = [C.c_void_p, C.c_int]
            self.lib.mtgl_dev_finish.argtypes = [C.c_void_p]
            self.lib.mtgl_dev_backend.restype = C.c_char_p
            vp = C.c_void_p
            for name, args in (("mtgl_batch_create", [C.POINTER(vp)]), ("mtgl_batch_begin", [vp]),
                               ("mtgl_batch_end", [vp, C.c_int]), ("mtgl_batch_launch", [vp]),
                               ("mtgl_batch_sync", [vp]), ("mtgl_batch_info", [vp, C.POINTER(MtglBatchInfo)]),
                               ("mtgl_batch_stats_get", [vp, C.POINTER(MtglDevStats)]),
                               ("mtgl_batch_stats_reset", [vp])):
                fn = getattr(self.lib, name)
                fn.argtypes = args
                fn.restype = C.c_int
            self.lib.mtgl_batch_destroy.argtypes = [vp]
            self.lib.mtgl_batch_destroy.restype = None
            self.lib.mtgl_texture_lod_mode.argtypes = [vp, C.c_uint32]
            self.lib.mtgl_texture_lod_mode.restype = None
        self._read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        self._read.restype = C.c_int

    def _wrap(self, fn, args):
        if "p" not in args:
            return fn

        def call(*a):
            conv = [(_ptr(v) if code == "p" else v) for code, v in zip(args, a)]
            return fn(*conv)
        call.__name__ = fn.__name__
        return call

    # ---- context helpers -------------------------------------------------
    def create_context(self, width: int, height: int) -> int:
        ctx = self.gl_create_context(width, height)
        if not ctx:
            raise RuntimeError(f"{self.kind}: gl_create_context({width}, {height}) failed")
        self.gl_make_current(ctx)
        return ctx

    def texture_lod_mode(self, ctx: int, mode: int) -> None:
        """MTGL_LOD_REFERENCE (0) or MTGL_LOD_FULL_CHAIN (1) for the context's
        later draws (extension; see include/mtgl_dev.h)."""
        self.lib.mtgl_texture_lod_mode(ctx, mode)

    def destroy_context(self, ctx: int) -> None:
        self.gl_make_current(None)
        self.gl_destroy_context(ctx)

    def read_framebuffer(self, ctx: int, width: int, height: int):
        color = np.empty((height, width), np.uint32)
        depth = np.empty((height, width), np.float32)
        stencil = np.empty((height, width), np.uint8)
        rc = self._read(ctx, color.ctypes.data, depth.ctypes.data, stencil.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"{self.kind}: framebuffer read failed ({rc})")
        return color, depth, stencil

    def device(self, ctx: int) -> int:
        return self.lib.mtgl_context_device(ctx)

    def stats(self, ctx: int) -> MtglDevStats:
        s = MtglDevStats()
        self.lib.mtgl_dev_stats_get(self.device(ctx), C.byref(s))
        return s

    def stats_reset(self, ctx: int) -> None:
        self.lib.mtgl_dev_stats_reset(self.device(ctx))

    def set_timing(self, ctx: int, on: bool, every: int = 1) -> None:
        """Time one fragment-stage launch in `every` with HIP events (on), or none."""
        self.lib.mtgl_dev_set_timing(self.device(ctx), max(1, int(every)) if on else 0)

    def finish(self, ctx: int) -> None:
        self.lib.mtgl_dev_finish(self.device(ctx))

    def batch(self) -> Batch:
        return Batch(self)

    def backend(self) -> str:
        if self.kind == "reference":
            return "reference:cpu"
        return self.lib.mtgl_dev_backend().decode()

    # ---- array helpers ---------------------------------------------------
    def keep(self, obj):
        """Hold a client array alive while the library may read it."""
        self._keep.append(obj)
        return obj

    def release(self) -> None:
        self._keep.clear()

    def fv(self, *vals) -> np.ndarray:
        return self.keep(np.asarray(vals, np.float32))
