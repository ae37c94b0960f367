"""Differentiable shading through the per-draw pipeline (include/mtgl_diff.h).

Views of one mesh are rendered with the ordinary GL path, each on its own
context, and captured (per-pixel visibility, vertex state, geometry).  Shading
is then recomputed from three parameter tensors -- vertex colours, the lights'
diffuse colours and an fp32 texture -- forward and adjoint, with the geometry
held fixed.  ``render`` is a ``torch.autograd.Function`` over the device
library; ``DiffViews.forward_np`` / ``backward_np`` drive any library with this
ABI on host arrays (the CPU oracle in tests).

With torch, import torch before the library is loaded: the HIP runtime torch
brings must be the process's first (loaded after libmtgl.so's, it reports no
GPU; loaded first, it serves both).

The workload (``diff_params``): a grid sphere of 2 g^2 triangles with radial
N(0, 0.05) displacement, per-vertex RGBA colours, a square RGBA texture, eight
directional lights fixed in eye space with random diffuse colours, and views on
a Fibonacci sphere looking at the origin.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .gl import ENUMS as E
from .scenes import SEED_BASE, sphere_grid

MAX_LIGHTS = 8


class MtglDiffInfo(C.Structure):
    """mtgl_diff_info_t of include/mtgl_diff.h."""
    _fields_ = [("views", C.c_uint32), ("width", C.c_int32), ("height", C.c_int32), ("vertices", C.c_uint32),
                ("tex_w", C.c_int32), ("tex_h", C.c_int32), ("work_floats", C.c_uint64)]


GPX_DTYPE = np.dtype([("tri", "<i4"), ("b", "<f4", (3,)), ("texel", "<u4", (4,)), ("fx", "<f4"), ("fy", "<f4")])


def _bind(lib):
    vp = C.c_void_p
    for name, args in (("mtgl_diff_create", [C.c_uint32, C.c_int32, C.c_int32, C.POINTER(vp)]),
                       ("mtgl_diff_add_view", [vp, vp]), ("mtgl_diff_arm", [vp]),
                       ("mtgl_diff_info", [vp, C.POINTER(MtglDiffInfo)]),
                       ("mtgl_diff_gbuffer", [vp, C.c_uint32, vp]),
                       ("mtgl_diff_forward", [vp, vp, vp, vp, vp, vp]),
                       ("mtgl_diff_backward", [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp])):
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mtgl_diff_destroy.argtypes = [vp]
    lib.mtgl_diff_destroy.restype = None


def diff_params(grid: int = 224, tex_size: int = 1024, views: int = 64, seed: int = SEED_BASE + 6):
    """Seeded parameters of the differentiable workload."""
    rng = np.random.default_rng(seed)
    verts, idx = sphere_grid(grid, seed)
    nv = verts.shape[0]
    vcol = np.ones((nv, 4), np.float32)
    vcol[:, :3] = rng.random((nv, 3), dtype=np.float32)
    tex = rng.random((tex_size, tex_size, 4), dtype=np.float32)
    d = rng.standard_normal((MAX_LIGHTS, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = np.abs(d[:, 2])            # lights on the viewer's side
    light_dir = d.astype(np.float32)
    light_rgb = rng.random((MAX_LIGHTS, 3), dtype=np.float32)
    k = np.arange(views) + 0.5
    y = 1.0 - 2.0 * k / views
    r = np.sqrt(1.0 - y * y)
    phi = math.pi * (3.0 - math.sqrt(5.0)) * k
    eyes = (3.0 * np.stack([r * np.cos(phi), y, r * np.sin(phi)], -1)).astype(np.float64)
    return {"verts": verts, "idx": idx, "vcol": vcol, "tex": tex, "light_dir": light_dir, "light_rgb": light_rgb,
            "eyes": eyes, "grid": grid}


def perturbed(params, seed: int, scale: float = 0.25):
    """The parameter set the targets are rendered from: vertex colours, light
    colours and texels moved by seeded uniform noise, clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    out = dict(params)
    for key in ("vcol", "light_rgb", "tex"):
        a = params[key]
        n = (rng.random(a.shape, dtype=np.float32) - 0.5) * (2.0 * scale)
        b = np.clip(a + n, 0.0, 1.0).astype(np.float32)
        if key == "vcol":
            b[:, 3] = a[:, 3]
        out[key] = b
    return out


def look_at(eye):
    """Column-major modelview looking from eye at the origin, y up."""
    e = np.asarray(eye, np.float64)
    f = -e / np.linalg.norm(e)
    up = np.array([0.0, 1.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ e
    return np.ascontiguousarray(m.T.reshape(-1).astype(np.float32))


def quantise(tex):
    """RGBA8 bytes of an fp32 texture and their exact fp32 values byte / 255."""
    b = np.clip(np.floor(tex * 255.0 + 0.5), 0, 255).astype(np.uint8)
    return b, (b.astype(np.float32) / np.float32(255.0))


def draw_view(g, params, view: int, W: int, H: int, tex8):
    """GL calls of one view on the current context: state, lights, texture,
    buffer objects, then one clear and one glDrawElements."""
    verts, idx = params["verts"], params["idx"]
    g.glViewport(0, 0, W, H)
    g.glClearColor(0.1, 0.1, 0.1, 1.0)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    a = W / H
    g.glFrustum(-0.1 * a, 0.1 * a, -0.1, 0.1, 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_LIGHTING"])
    for i in range(MAX_LIGHTS):
        li = E["GL_LIGHT0"] + i
        g.glEnable(li)
        d = params["light_dir"][i]
        g.glLightfv(li, E["GL_POSITION"], g.fv(float(d[0]), float(d[1]), float(d[2]), 0.0))
        c = params["light_rgb"][i]
        g.glLightfv(li, E["GL_DIFFUSE"], g.fv(float(c[0]), float(c[1]), float(c[2]), 1.0))
        g.glLightfv(li, E["GL_SPECULAR"], g.fv(0.0, 0.0, 0.0, 1.0))
    g.glLoadMatrixf(g.keep(look_at(params["eyes"][view])))
    g.glEnable(E["GL_COLOR_MATERIAL"])
    g.glColorMaterial(E["GL_FRONT_AND_BACK"], E["GL_AMBIENT_AND_DIFFUSE"])
    g.glShadeModel(E["GL_SMOOTH"])
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glHint(E["GL_PERSPECTIVE_CORRECTION_HINT"], E["GL_NICEST"])
    ts = tex8.shape[0]
    ids = np.zeros(1, np.uint32)
    g.glGenTextures(1, ids)
    g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[0]))
    g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], ts, ts, 0, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"],
                   g.keep(np.ascontiguousarray(tex8)))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_S"], E["GL_REPEAT"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_T"], E["GL_REPEAT"])
    g.glTexEnvi(E["GL_TEXTURE_ENV"], E["GL_TEXTURE_ENV_MODE"], E["GL_MODULATE"])
    g.glEnable(E["GL_TEXTURE_2D"])
    # one interleaved buffer object: the reference resolves every array
    # against the buffer bound at draw time (src/gl_api.c:1745-1755)
    inter = np.ascontiguousarray(np.concatenate([verts, np.asarray(params["vcol"], np.float32)], 1))
    bufs = np.zeros(2, np.uint32)
    g.glGenBuffers(2, bufs)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(bufs[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], inter.nbytes, g.keep(inter), E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glEnableClientState(E["GL_TEXTURE_COORD_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 48, 0)
    g.glNormalPointer(E["GL_FLOAT"], 48, 12)
    g.glTexCoordPointer(2, E["GL_FLOAT"], 48, 24)
    g.glColorPointer(4, E["GL_FLOAT"], 48, 32)
    g.glBindBuffer(E["GL_ELEMENT_ARRAY_BUFFER"], int(bufs[1]))
    g.glBufferData(E["GL_ELEMENT_ARRAY_BUFFER"], idx.nbytes, g.keep(idx), E["GL_STATIC_DRAW"])
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
    g.glDrawElements(E["GL_TRIANGLES"], len(idx), E["GL_UNSIGNED_INT"], 0)


class DiffViews:
    """Views of params rendered on library g and captured into one mtgl_diff.
    The ordinary renders (RGBA8, with the quantised texture) are kept in
    ``frames`` when keep_frames is set."""

    def __init__(self, g, params, W: int, H: int, keep_frames: bool = False):
        self.g = g
        _bind(g.lib)
        self.W, self.H = W, H
        self.nviews = len(params["eyes"])
        self.nvert = params["verts"].shape[0]
        tex8, _ = quantise(params["tex"])
        ts = tex8.shape[0]
        h = C.c_void_p()
        rc = g.lib.mtgl_diff_create(self.nvert, ts, ts, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"{g.kind}: mtgl_diff_create failed ({rc})")
        self.h = h.value
        self.frames = []
        try:
            for v in range(self.nviews):
                ctx = g.create_context(W, H)
                try:
                    rc = g.lib.mtgl_diff_arm(g.device(ctx))
                    if rc != 0:
                        raise RuntimeError(f"{g.kind}: mtgl_diff_arm failed ({rc})")
                    draw_view(g, params, v, W, H, tex8)
                    if keep_frames:
                        self.frames.append(g.read_framebuffer(ctx, W, H)[0])
                    rc = g.lib.mtgl_diff_add_view(self.h, g.device(ctx))
                    if rc != 0:
                        raise RuntimeError(f"{g.kind}: mtgl_diff_add_view (view {v}) failed ({rc})")
                finally:
                    g.destroy_context(ctx)
                    g.release()
        except Exception:
            self.close()
            raise
        info = MtglDiffInfo()
        g.lib.mtgl_diff_info(self.h, C.byref(info))
        self.info = info

    # ---- argument checks: the kernels read and write these exact layouts
    def shapes(self):
        """Expected shapes of vcol, light_rgb, tex and of an image / its
        gradient: [vertices][4], [MAX_LIGHTS][3], [tex_h][tex_w][4] and
        [views][H][W][4], all float32."""
        i = self.info
        return {"vcol": (int(i.vertices), 4), "light_rgb": (MAX_LIGHTS, 3), "tex": (int(i.tex_h), int(i.tex_w), 4),
                "grad": (self.nviews, self.H, self.W, 4)}

    def check(self, **arrays):
        """Raise ValueError unless every named array has its expected shape
        (see shapes()); dtypes are checked by the callers."""
        want = self.shapes()
        for name, a in arrays.items():
            if tuple(a.shape) != want[name]:
                raise ValueError(f"{name} has shape {tuple(a.shape)}, expected {want[name]}")

    def gbuffer(self, view: int) -> np.ndarray:
        out = np.empty(self.W * self.H, GPX_DTYPE)
        rc = self.g.lib.mtgl_diff_gbuffer(self.h, view, out.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"mtgl_diff_gbuffer failed ({rc})")
        return out.reshape(self.H, self.W)

    # ---- host arrays (the CPU oracle)
    def forward_np(self, vcol, light_rgb, tex):
        a = [np.ascontiguousarray(x, np.float32) for x in (vcol, light_rgb, tex)]
        self.check(vcol=a[0], light_rgb=a[1], tex=a[2])
        img = np.empty((self.nviews, self.H, self.W, 4), np.float32)
        rc = self.g.lib.mtgl_diff_forward(self.h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                          img.ctypes.data, None)
        if rc != 0:
            raise RuntimeError(f"mtgl_diff_forward failed ({rc})")
        return img

    def backward_np(self, vcol, light_rgb, tex, grad):
        a = [np.ascontiguousarray(x, np.float32) for x in (vcol, light_rgb, tex, grad)]
        self.check(vcol=a[0], light_rgb=a[1], tex=a[2], grad=a[3])
        gv = np.empty_like(a[0])
        gl = np.empty((MAX_LIGHTS, 3), np.float32)
        gt = np.empty_like(a[2])
        work = np.empty(max(1, int(self.info.work_floats)), np.float32)
        rc = self.g.lib.mtgl_diff_backward(self.h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                           a[3].ctypes.data, gv.ctypes.data, gl.ctypes.data, gt.ctypes.data,
                                           work.ctypes.data, None)
        if rc != 0:
            raise RuntimeError(f"mtgl_diff_backward failed ({rc})")
        return gv, gl, gt

    def close(self):
        if self.h:
            self.g.lib.mtgl_diff_destroy(self.h)
            self.h = None


def _torch_fn():
    import torch

    class Render(torch.autograd.Function):
        """images = shade(vcol, light_rgb, tex) over the captured views; the
        backward pass is the device adjoint (mtgl_diff_backward)."""

        @staticmethod
        def forward(ctx, vcol, light_rgb, tex, views):
            for t in (vcol, light_rgb, tex):
                if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                    raise ValueError("vcol, light_rgb and tex must be contiguous float32 device tensors")
                if t.device != vcol.device:
                    raise ValueError("vcol, light_rgb and tex must be on one device")
            views.check(vcol=vcol, light_rgb=light_rgb, tex=tex)
            img = torch.empty((views.nviews, views.H, views.W, 4), device=vcol.device, dtype=torch.float32)
            st = torch.cuda.current_stream(vcol.device).cuda_stream
            rc = views.g.lib.mtgl_diff_forward(views.h, vcol.data_ptr(), light_rgb.data_ptr(), tex.data_ptr(),
                                               img.data_ptr(), st)
            if rc != 0:
                raise RuntimeError(f"mtgl_diff_forward failed ({rc})")
            ctx.save_for_backward(vcol, light_rgb, tex)
            ctx.views = views
            return img

        @staticmethod
        def backward(ctx, grad):
            vcol, light_rgb, tex = ctx.saved_tensors
            views = ctx.views
            grad = grad.to(device=vcol.device, dtype=torch.float32).contiguous()
            views.check(grad=grad)
            gv = torch.empty_like(vcol)
            gl = torch.empty_like(light_rgb)
            gt = torch.empty_like(tex)
            work = torch.empty(max(1, int(views.info.work_floats)), device=vcol.device, dtype=torch.float32)
            st = torch.cuda.current_stream(vcol.device).cuda_stream
            rc = views.g.lib.mtgl_diff_backward(views.h, vcol.data_ptr(), light_rgb.data_ptr(), tex.data_ptr(),
                                                grad.data_ptr(), gv.data_ptr(), gl.data_ptr(), gt.data_ptr(),
                                                work.data_ptr(), st)
            if rc != 0:
                raise RuntimeError(f"mtgl_diff_backward failed ({rc})")
            return gv, gl, gt, None

    return Render


_RENDER = None


def render(vcol, light_rgb, tex, views: DiffViews):
    """Differentiable render of every captured view (torch tensors on the
    device of the library's contexts)."""
    global _RENDER
    if _RENDER is None:
        _RENDER = _torch_fn()
    return _RENDER.apply(vcol, light_rgb, tex, views)
