"""Seeded scenes that drive a GLLibrary through the GL 1.5 API.

Every scene is a function ``scene(g, W, H, seed)`` issuing GL calls on a bound
library ``g`` whose context is current.  ``render()`` creates the context,
clears every plane, runs the scene and returns (colour, depth, stencil).

The workload configurations are the ones the benchmark and the parity tests
use:

* ``icosphere``  - 1,280-triangle icosphere, per-vertex colours through
  GL_COLOR_MATERIAL, one directional light, Gouraud, depth test, 640x480,
  glDrawArrays.
* ``textured_sphere`` - 128x128 lat-long grid sphere (32,768 triangles) with
  radial N(0, 0.05) displacement, 256x256 value-noise texture with
  GL_LINEAR_MIPMAP_LINEAR / GL_LINEAR, perspective-correct, linear fog, one
  light, 1920x1080, glDrawElements from buffer objects.
* ``overdraw`` - 200k triangles with log-uniform areas of 1..1e4 px, blended
  SRC_ALPHA / ONE_MINUS_SRC_ALPHA, depth LESS, stencil GREATER 8 / INCR,
  back-face culling, 3840x2160.
* ``batch`` - 512 independent scenes, one context each: square framebuffers
  with log-uniform sides 64..1024, grid spheres of log-uniform 100..50k
  triangles, every other one textured (power-of-two sides 64..512,
  GL_LINEAR_MIPMAP_LINEAR), one clear and one glDrawElements per frame.

The remaining scenes are small feature scenes that together reach every branch
of the per-vertex, clipping and per-fragment paths (reference src/raster.c,
src/lighting.h, src/textures.c, src/clipping.h).
"""
from __future__ import annotations

import math

import numpy as np

from .gl import ENUMS as E

SEED_BASE = 0x4D54474C00          # seeds 0x4D54474C01.. for the configurations


def _f32(*v):
    return np.asarray(v, np.float32)


def _rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------------------
# geometry and texture generators
# ---------------------------------------------------------------------------

def icosphere(subdiv: int):
    """Unit icosphere: (positions (V,3) float64, triangles (T,3) int)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        new = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new
    return np.array(verts), np.array(faces, np.int64)


def value_noise_rgba(size: int, seed: int, octaves: int = 4, lattice: int = 16) -> np.ndarray:
    """size x size RGBA8 value noise: `octaves` octaves of a seeded lattice^2
    grid per channel, smoothstep interpolation, summed with halving weights and
    quantised to bytes."""
    rng = _rng(seed)
    out = np.zeros((size, size, 4), np.float64)
    total = 0.0
    for o in range(octaves):
        cells = lattice * (2 ** o)
        grid = rng.random((cells + 1, cells + 1, 4))
        grid[cells, :] = grid[0, :]
        grid[:, cells] = grid[:, 0]
        coord = (np.arange(size) + 0.5) * cells / size
        i0 = np.floor(coord).astype(np.int64)
        f = coord - i0
        s = f * f * (3 - 2 * f)
        gy0, gy1 = i0[:, None], i0[:, None] + 1
        gx0, gx1 = i0[None, :], i0[None, :] + 1
        sx, sy = s[None, :, None], s[:, None, None]
        top = grid[gy0, gx0] * (1 - sx) + grid[gy0, gx1] * sx
        bot = grid[gy1, gx0] * (1 - sx) + grid[gy1, gx1] * sx
        w = 0.5 ** o
        out += w * (top * (1 - sy) + bot * sy)
        total += w
    return np.clip(np.floor(out / total * 256.0), 0, 255).astype(np.uint8)


def random_rgba(w: int, h: int, seed: int) -> np.ndarray:
    return _rng(seed).integers(0, 256, size=(h, w, 4), dtype=np.uint8)


# ---------------------------------------------------------------------------
# runner
# ---------------------------------------------------------------------------

def render(g, scene, W: int, H: int, seed: int = 0, clear=(0.0, 0.0, 0.0, 1.0), keep_context=False):
    ctx = g.create_context(W, H)
    try:
        g.glClearColor(*clear)
        g.glClearDepth(1.0)
        g.glClearStencil(0)
        g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
        scene(g, W, H, seed)
        out = g.read_framebuffer(ctx, W, H)
        err = g.glGetError()
    finally:
        if not keep_context:
            g.destroy_context(ctx)
            g.release()
    if keep_context:
        return out, err, ctx
    return out, err


def _perspective(g, W, H, near=0.1, far=100.0, half_h=0.1, dist=-3.0):
    aspect = W / H
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glFrustum(-half_h * aspect, half_h * aspect, -half_h, half_h, near, far)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    if dist:
        g.glTranslatef(0.0, 0.0, dist)


# ---------------------------------------------------------------------------
# workload configurations
# ---------------------------------------------------------------------------

def setup_icosphere(g, W=640, H=480, seed=SEED_BASE + 1):
    pos, faces = icosphere(3)
    rng = _rng(seed)
    colors = rng.random((len(pos), 4)).astype(np.float32)
    tri = faces.reshape(-1)
    vp = g.keep(np.ascontiguousarray(pos[tri].astype(np.float32)))
    vn = g.keep(vp.copy())
    vc = g.keep(np.ascontiguousarray(colors[tri]))
    g.glViewport(0, 0, W, H)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glFrustum(-0.1 * 4.0 / 3.0, 0.1 * 4.0 / 3.0, -0.1, 0.1, 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.fv(1, 1, 1, 0))
    g.glEnable(E["GL_COLOR_MATERIAL"])
    g.glColorMaterial(E["GL_FRONT_AND_BACK"], E["GL_AMBIENT_AND_DIFFUSE"])
    g.glShadeModel(E["GL_SMOOTH"])
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glDepthFunc(E["GL_LESS"])
    g.glTranslatef(0.0, 0.0, -3.0)
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 0, vp)
    g.glNormalPointer(E["GL_FLOAT"], 0, vn)
    g.glColorPointer(4, E["GL_FLOAT"], 0, vc)
    return len(tri)


def draw_icosphere(g, n):
    g.glDrawArrays(E["GL_TRIANGLES"], 0, n)


def scene_icosphere(g, W, H, seed=SEED_BASE + 1):
    draw_icosphere(g, setup_icosphere(g, W, H, seed))


def sphere_grid(n: int, seed: int, sigma: float = 0.05):
    """(n+1)^2 lat-long vertices with per-unique-vertex radial N(0, sigma) noise
    (seam column and poles shared) and 2 n^2 triangles."""
    rng = _rng(seed)
    noise = rng.standard_normal((n + 1, n)) * sigma
    noise[0, :] = noise[0, 0]
    noise[n, :] = noise[n, 0]
    i = np.arange(n + 1)[:, None]
    j = np.arange(n + 1)[None, :]
    theta = np.pi * i / n
    phi = 2 * np.pi * j / n
    unit = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta) * np.ones_like(phi),
                     np.sin(theta) * np.sin(phi)], -1)
    r = 1.0 + noise[:, np.arange(n + 1) % n]
    posn = unit * r[..., None]
    uv = np.stack([np.broadcast_to(j / n, (n + 1, n + 1)), np.broadcast_to(i / n, (n + 1, n + 1))], -1)
    verts = np.concatenate([posn, unit, uv], -1).reshape(-1, 8).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v00 = ii * (n + 1) + jj
    v01 = v00 + 1
    v10 = v00 + (n + 1)
    v11 = v10 + 1
    idx = np.stack([v00, v10, v11, v00, v11, v01], -1).reshape(-1).astype(np.uint32)
    return np.ascontiguousarray(verts), np.ascontiguousarray(idx)


def setup_textured_sphere(g, W=1920, H=1080, seed=SEED_BASE + 2, grid=128, tex_size=256):
    verts, idx = sphere_grid(grid, seed)
    tex = value_noise_rgba(tex_size, SEED_BASE + 3)
    g.glViewport(0, 0, W, H)
    aspect = 16.0 / 9.0
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glFrustum(-0.1 * aspect, 0.1 * aspect, -0.1, 0.1, 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.fv(1, 1, 1, 0))
    g.glTranslatef(0.0, 0.0, -2.0)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glHint(E["GL_PERSPECTIVE_CORRECTION_HINT"], E["GL_NICEST"])
    g.glEnable(E["GL_FOG"])
    g.glFogi(E["GL_FOG_MODE"], E["GL_LINEAR"])
    g.glFogf(E["GL_FOG_START"], 1.0)
    g.glFogf(E["GL_FOG_END"], 3.0)
    ids = np.zeros(1, np.uint32)
    g.glGenTextures(1, ids)
    g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[0]))
    g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], tex_size, tex_size, 0, E["GL_RGBA"],
                   E["GL_UNSIGNED_BYTE"], g.keep(np.ascontiguousarray(tex)))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR_MIPMAP_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_S"], E["GL_REPEAT"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_T"], E["GL_REPEAT"])
    g.glEnable(E["GL_TEXTURE_2D"])
    bufs = np.zeros(2, np.uint32)
    g.glGenBuffers(2, bufs)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(bufs[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], verts.nbytes, g.keep(verts), E["GL_STATIC_DRAW"])
    g.glBindBuffer(E["GL_ELEMENT_ARRAY_BUFFER"], int(bufs[1]))
    g.glBufferData(E["GL_ELEMENT_ARRAY_BUFFER"], idx.nbytes, g.keep(idx), E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glEnableClientState(E["GL_TEXTURE_COORD_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 32, 0)
    g.glNormalPointer(E["GL_FLOAT"], 32, 12)
    g.glTexCoordPointer(2, E["GL_FLOAT"], 32, 24)
    return len(idx)


def draw_textured_sphere(g, n):
    g.glDrawElements(E["GL_TRIANGLES"], n, E["GL_UNSIGNED_INT"], 0)


def scene_textured_sphere(g, W, H, seed=SEED_BASE + 2):
    draw_textured_sphere(g, setup_textured_sphere(g, W, H, seed))


def setup_libm_spheres(g, W=1920, H=1080, seed=SEED_BASE + 7, grid=96):
    """Two grid spheres at 1920x1080 whose every pixel goes through the
    libm-dependent steps of the fragment path: GL_PHONG lighting per pixel
    (Blinn-Phong specular, powf, src/lighting.h:126-132) from a positional spot
    light (spot exponent, powf, :103-110) and a directional light, one sphere
    under GL_EXP fog and one under GL_EXP2 fog (expf, src/raster.c:686-691).
    Returns the index count of one sphere."""
    verts, idx = sphere_grid(grid, seed)
    g.glViewport(0, 0, W, H)
    aspect = W / H
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glFrustum(-0.1 * aspect, 0.1 * aspect, -0.1, 0.1, 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_LIGHTING"])
    g.glShadeModel(E["GL_PHONG"])
    g.glEnable(E["GL_LIGHT0"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.fv(0.3, 0.5, 1.0, 0.0))
    g.glLightfv(E["GL_LIGHT0"], E["GL_DIFFUSE"], g.fv(0.5, 0.5, 0.45, 1.0))
    g.glLightfv(E["GL_LIGHT0"], E["GL_SPECULAR"], g.fv(0.6, 0.6, 0.6, 1.0))
    g.glEnable(E["GL_LIGHT1"])
    g.glLightfv(E["GL_LIGHT1"], E["GL_POSITION"], g.fv(0.0, 1.5, 0.5, 1.0))
    g.glLightfv(E["GL_LIGHT1"], E["GL_SPOT_DIRECTION"], g.fv(0.0, -1.0, -1.2))
    g.glLightf(E["GL_LIGHT1"], E["GL_SPOT_CUTOFF"], 35.0)
    g.glLightf(E["GL_LIGHT1"], E["GL_SPOT_EXPONENT"], 7.5)
    g.glLightf(E["GL_LIGHT1"], E["GL_LINEAR_ATTENUATION"], 0.2)
    g.glLightfv(E["GL_LIGHT1"], E["GL_DIFFUSE"], g.fv(0.9, 0.7, 0.4, 1.0))
    g.glLightfv(E["GL_LIGHT1"], E["GL_SPECULAR"], g.fv(1.0, 1.0, 1.0, 1.0))
    g.glLightModeli(E["GL_LIGHT_MODEL_LOCAL_VIEWER"], 1)
    g.glMaterialfv(E["GL_FRONT_AND_BACK"], E["GL_SPECULAR"], g.fv(0.8, 0.8, 0.8, 1.0))
    g.glMaterialf(E["GL_FRONT_AND_BACK"], E["GL_SHININESS"], 37.3)
    g.glMaterialfv(E["GL_FRONT_AND_BACK"], E["GL_DIFFUSE"], g.fv(0.7, 0.55, 0.4, 1.0))
    g.glEnable(E["GL_FOG"])
    g.glFogfv(E["GL_FOG_COLOR"], g.fv(0.3, 0.35, 0.4, 1.0))
    bufs = np.zeros(2, np.uint32)
    g.glGenBuffers(2, bufs)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(bufs[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], verts.nbytes, g.keep(verts), E["GL_STATIC_DRAW"])
    g.glBindBuffer(E["GL_ELEMENT_ARRAY_BUFFER"], int(bufs[1]))
    g.glBufferData(E["GL_ELEMENT_ARRAY_BUFFER"], idx.nbytes, g.keep(idx), E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 32, 0)
    g.glNormalPointer(E["GL_FLOAT"], 32, 12)
    return len(idx)


def draw_libm_spheres(g, n):
    for k, (mode, dens) in enumerate((("GL_EXP", 0.45), ("GL_EXP2", 0.38))):
        g.glFogi(E["GL_FOG_MODE"], E[mode])
        g.glFogf(E["GL_FOG_DENSITY"], dens)
        g.glLoadIdentity()
        g.glTranslatef(-0.75 + 1.5 * k, 0.0, -2.6 - 0.4 * k)
        g.glRotatef(30.0 + 50.0 * k, 0.3, 1.0, 0.1)
        g.glDrawElements(E["GL_TRIANGLES"], n, E["GL_UNSIGNED_INT"], 0)


def overdraw_geometry(W: int, H: int, ntri: int, seed: int, max_log_area: float = 4.0):
    """Triangles with area 10^U(0, max_log_area) px, centroid ~ U(screen), random
    rotation and aspect in [1, 4], constant eye depth U(0.1, 100), RGB U[0,1],
    alpha U(0.2, 1), winding flipped for half of them."""
    rng = _rng(seed)
    area = 10.0 ** rng.uniform(0.0, max_log_area, ntri)
    cx = rng.uniform(0, W, ntri)
    cy = rng.uniform(0, H, ntri)
    rot = rng.uniform(0, 2 * np.pi, ntri)
    aspect = rng.uniform(1.0, 4.0, ntri)
    depth = rng.uniform(0.1, 100.0, ntri)
    rgb = rng.random((ntri, 3))
    alpha = rng.uniform(0.2, 1.0, ntri)
    flip = rng.random(ntri) < 0.5
    ang = np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])
    bx = np.cos(ang)[None, :] * aspect[:, None]
    by = np.sin(ang)[None, :]
    base_area = 0.5 * np.abs((bx[:, 1] - bx[:, 0]) * (by[:, 2] - by[:, 0]) - (bx[:, 2] - bx[:, 0]) * (by[:, 1] - by[:, 0]))
    s = np.sqrt(area / base_area)
    rx = (bx * np.cos(rot)[:, None] - by * np.sin(rot)[:, None]) * s[:, None] + cx[:, None]
    ry = (bx * np.sin(rot)[:, None] + by * np.cos(rot)[:, None]) * s[:, None] + cy[:, None]
    rx[flip, 1], rx[flip, 2] = rx[flip, 2].copy(), rx[flip, 1].copy()
    ry[flip, 1], ry[flip, 2] = ry[flip, 2].copy(), ry[flip, 1].copy()
    pos = np.stack([rx, ry, np.broadcast_to(-depth[:, None], rx.shape)], -1).reshape(-1, 3).astype(np.float32)
    col = np.concatenate([rgb, alpha[:, None]], 1)
    col = np.repeat(col, 3, axis=0).astype(np.float32)
    return np.ascontiguousarray(pos), np.ascontiguousarray(col)


def setup_overdraw(g, W=3840, H=2160, seed=SEED_BASE + 4, ntri=200_000, max_log_area=4.0):
    pos, col = overdraw_geometry(W, H, ntri, seed, max_log_area)
    g.glViewport(0, 0, W, H)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), 0.1, 100.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glEnable(E["GL_CULL_FACE"])
    g.glCullFace(E["GL_BACK"])
    g.glEnable(E["GL_BLEND"])
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glDepthFunc(E["GL_LESS"])
    g.glEnable(E["GL_STENCIL_TEST"])
    g.glStencilFunc(E["GL_GREATER"], 8, 0xFF)
    g.glStencilOp(E["GL_KEEP"], E["GL_KEEP"], E["GL_INCR"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 0, g.keep(pos))
    g.glColorPointer(4, E["GL_FLOAT"], 0, g.keep(col))
    return len(pos)


def draw_overdraw(g, n):
    g.glDrawArrays(E["GL_TRIANGLES"], 0, n)


def scene_overdraw(g, W, H, seed=SEED_BASE + 4, ntri=200_000):
    draw_overdraw(g, setup_overdraw(g, W, H, seed, ntri))


# ---------------------------------------------------------------------------
# feature scenes
# ---------------------------------------------------------------------------

def _random_tris_immediate(g, rng, n, spread=3.0, alpha=(0.0, 1.0), tex=False, normals=False):
    for _ in range(n):
        for _k in range(3):
            c = rng.random(4)
            c[3] = rng.uniform(*alpha)
            g.glColor4f(*c.astype(np.float32))
            if tex:
                g.glTexCoord2f(*rng.uniform(-1.0, 2.0, 2).astype(np.float32))
            if normals:
                g.glNormal3f(*rng.normal(size=3).astype(np.float32))
            g.glVertex3f(*((rng.random(3) - 0.5) * spread).astype(np.float32))


def scene_depth_funcs(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    funcs = ["GL_NEVER", "GL_LESS", "GL_EQUAL", "GL_LEQUAL", "GL_GREATER", "GL_NOTEQUAL", "GL_GEQUAL", "GL_ALWAYS"]
    for k, fn in enumerate(funcs):
        g.glDepthFunc(E[fn])
        g.glDepthMask(bool(k % 3 != 2))
        if k == 4:
            g.glDepthRange(0.2, 0.7)
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 6)
        g.glEnd()


def scene_blend(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_BLEND"])
    srcs = ["GL_ZERO", "GL_ONE", "GL_SRC_COLOR", "GL_ONE_MINUS_SRC_COLOR", "GL_DST_COLOR",
            "GL_ONE_MINUS_DST_COLOR", "GL_SRC_ALPHA", "GL_ONE_MINUS_SRC_ALPHA", "GL_DST_ALPHA",
            "GL_ONE_MINUS_DST_ALPHA", "GL_SRC_ALPHA_SATURATE", "GL_CONSTANT_COLOR", "GL_ONE_MINUS_CONSTANT_ALPHA"]
    dsts = [s for s in srcs if s != "GL_SRC_ALPHA_SATURATE"]
    for k in range(24):
        g.glBlendFunc(E[srcs[rng.integers(len(srcs))]], E[dsts[rng.integers(len(dsts))]])
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 3)
        g.glEnd()


def scene_stencil(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_STENCIL_TEST"])
    g.glEnable(E["GL_DEPTH_TEST"])
    funcs = ["GL_NEVER", "GL_LESS", "GL_EQUAL", "GL_LEQUAL", "GL_GREATER", "GL_NOTEQUAL", "GL_GEQUAL", "GL_ALWAYS"]
    ops = ["GL_KEEP", "GL_ZERO", "GL_REPLACE", "GL_INCR", "GL_INCR_WRAP", "GL_DECR", "GL_DECR_WRAP", "GL_INVERT"]
    for k in range(30):
        g.glStencilFunc(E[funcs[rng.integers(8)]], int(rng.integers(-3, 300)), int(rng.choice([0xFF, 0x0F, 0xF0, 0x7])))
        g.glStencilOp(E[ops[rng.integers(8)]], E[ops[rng.integers(8)]], E[ops[rng.integers(8)]])
        g.glStencilMask(int(rng.choice([0xFF, 0xFFFFFFFF, 0x3C, 0x01])))
        g.glDepthFunc(E[funcs[1 + rng.integers(7)]])
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 2)
        g.glEnd()


def _upload_texture(g, data: np.ndarray, fmt="GL_RGBA"):
    ids = np.zeros(1, np.uint32)
    g.glGenTextures(1, ids)
    g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[0]))
    h, w = data.shape[:2]
    g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E[fmt], w, h, 0, E[fmt], E["GL_UNSIGNED_BYTE"],
                   g.keep(np.ascontiguousarray(data)))
    return int(ids[0])


def scene_textures(g, W, H, seed):
    """Every min/mag filter, wrap mode and texture-environment mode, both
    perspective hints, odd texture sizes (including < 2x2 without a level 1),
    a texture matrix with w != 1, and texel alpha with the alpha test."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_TEXTURE_2D"])
    g.glEnable(E["GL_DEPTH_TEST"])
    texes = [_upload_texture(g, random_rgba(16, 16, seed)), _upload_texture(g, random_rgba(7, 5, seed + 1)),
             _upload_texture(g, random_rgba(1, 9, seed + 2)),
             _upload_texture(g, rng.integers(0, 256, (8, 8, 3), dtype=np.uint8), "GL_RGB"),
             _upload_texture(g, rng.integers(0, 256, (4, 12), dtype=np.uint8), "GL_LUMINANCE"),
             _upload_texture(g, rng.integers(0, 256, (6, 6, 2), dtype=np.uint8), "GL_LUMINANCE_ALPHA")]
    mins = ["GL_NEAREST", "GL_LINEAR", "GL_NEAREST_MIPMAP_NEAREST", "GL_LINEAR_MIPMAP_NEAREST",
            "GL_NEAREST_MIPMAP_LINEAR", "GL_LINEAR_MIPMAP_LINEAR"]
    wraps = ["GL_REPEAT", "GL_CLAMP", "GL_CLAMP_TO_EDGE"]
    envs = ["GL_MODULATE", "GL_DECAL", "GL_REPLACE", "GL_BLEND", "GL_ADD"]
    g.glTexEnvfv(E["GL_TEXTURE_ENV"], E["GL_TEXTURE_ENV_COLOR"], g.fv(0.3, 0.6, 0.9, 0.5))
    for k in range(36):
        g.glBindTexture(E["GL_TEXTURE_2D"], texes[k % len(texes)])
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E[mins[k % 6]])
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E[["GL_NEAREST", "GL_LINEAR"][(k // 6) % 2]])
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_S"], E[wraps[k % 3]])
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_WRAP_T"], E[wraps[(k // 3) % 3]])
        g.glTexEnvi(E["GL_TEXTURE_ENV"], E["GL_TEXTURE_ENV_MODE"], E[envs[k % 5]])
        g.glHint(E["GL_PERSPECTIVE_CORRECTION_HINT"], E[["GL_NICEST", "GL_FASTEST", "GL_DONT_CARE"][k % 3]])
        if k % 7 == 3:
            g.glEnable(E["GL_ALPHA_TEST"])
            g.glAlphaFunc(E[["GL_GREATER", "GL_LESS", "GL_GEQUAL", "GL_NOTEQUAL"][k % 4]], float(rng.random()))
        else:
            g.glDisable(E["GL_ALPHA_TEST"])
        g.glMatrixMode(E["GL_TEXTURE"])
        g.glLoadIdentity()
        if k % 5 == 1:
            g.glTranslatef(0.25, -0.5, 0.0)
            g.glScalef(3.0, 0.5, 1.0)
        if k % 9 == 4:
            g.glMultMatrixf(g.fv(1, 0, 0, 0.2, 0, 1, 0, 0.1, 0, 0, 1, 0, 0, 0, 0, 1.5))
        g.glMatrixMode(E["GL_MODELVIEW"])
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 2, spread=(0.6 if k % 4 == 0 else 3.0), tex=True)
        g.glEnd()


def scene_fog(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H, dist=0.0)
    g.glEnable(E["GL_FOG"])
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glFogfv(E["GL_FOG_COLOR"], g.fv(0.7, 0.2, 0.4, 0.6))
    for k, mode in enumerate(["GL_LINEAR", "GL_EXP", "GL_EXP2", "GL_LINEAR"]):
        g.glFogi(E["GL_FOG_MODE"], E[mode])
        g.glFogf(E["GL_FOG_DENSITY"], float(rng.uniform(0.05, 0.6)))
        g.glFogf(E["GL_FOG_START"], 1.0)
        g.glFogf(E["GL_FOG_END"], 1.0 if k == 3 else 6.0)
        g.glBegin(E["GL_TRIANGLES"])
        for _ in range(5):
            for _v in range(3):
                g.glColor4f(*rng.random(4).astype(np.float32))
                p = rng.uniform(-2, 2, 3)
                p[2] = -rng.uniform(0.5, 9.0)
                g.glVertex3f(*p.astype(np.float32))
        g.glEnd()


def scene_lighting(g, W, H, seed):
    """Directional, positional (attenuated) and spot lights, specular with a
    local viewer, NORMALIZE, every colour-material mode, two-sided lighting,
    flat and Phong shading."""
    rng = _rng(seed)
    _perspective(g, W, H, dist=0.0)
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glLightModelfv(E["GL_LIGHT_MODEL_AMBIENT"], g.fv(0.1, 0.15, 0.2, 1.0))
    for li in range(4):
        L = E["GL_LIGHT0"] + li
        g.glEnable(L)
        g.glLightfv(L, E["GL_DIFFUSE"], g.fv(*rng.random(4)))
        g.glLightfv(L, E["GL_SPECULAR"], g.fv(*rng.random(4)))
        g.glLightfv(L, E["GL_AMBIENT"], g.fv(*(0.3 * rng.random(4))))
        w = 0.0 if li == 0 else 1.0
        g.glLightfv(L, E["GL_POSITION"], g.fv(*rng.uniform(-3, 3, 3), w))
        if li >= 1:
            g.glLightf(L, E["GL_CONSTANT_ATTENUATION"], 0.5)
            g.glLightf(L, E["GL_LINEAR_ATTENUATION"], 0.2)
            g.glLightf(L, E["GL_QUADRATIC_ATTENUATION"], 0.05)
        if li == 2:
            g.glLightfv(L, E["GL_SPOT_DIRECTION"], g.fv(0.0, 0.0, -1.0))
            g.glLightf(L, E["GL_SPOT_CUTOFF"], 40.0)
            g.glLightf(L, E["GL_SPOT_EXPONENT"], 3.0)
    g.glMaterialfv(E["GL_FRONT_AND_BACK"], E["GL_SPECULAR"], g.fv(0.8, 0.8, 0.8, 1.0))
    g.glMaterialf(E["GL_FRONT"], E["GL_SHININESS"], 20.0)
    g.glMaterialfv(E["GL_BACK"], E["GL_DIFFUSE"], g.fv(0.2, 0.9, 0.3, 0.7))
    g.glTranslatef(0.0, 0.0, -4.0)
    modes = ["GL_AMBIENT", "GL_DIFFUSE", "GL_SPECULAR", "GL_EMISSION", "GL_AMBIENT_AND_DIFFUSE"]
    faces = ["GL_FRONT", "GL_BACK", "GL_FRONT_AND_BACK"]
    for k in range(12):
        g.glShadeModel(E[["GL_SMOOTH", "GL_FLAT", "GL_PHONG"][k % 3]])
        g.glLightModeli(E["GL_LIGHT_MODEL_TWO_SIDE"], int(k % 4 >= 2))
        g.glLightModeli(E["GL_LIGHT_MODEL_LOCAL_VIEWER"], int(k % 5 == 1))
        if k % 2:
            g.glEnable(E["GL_NORMALIZE"])
        else:
            g.glDisable(E["GL_NORMALIZE"])
        if k % 3 != 2:
            g.glEnable(E["GL_COLOR_MATERIAL"])
            g.glColorMaterial(E[faces[k % 3]], E[modes[k % 5]])
        else:
            g.glDisable(E["GL_COLOR_MATERIAL"])
        g.glRotatef(float(rng.uniform(0, 360)), *rng.normal(size=3).astype(np.float32))
        if k == 7:
            g.glScalef(1.5, 0.5, 2.0)
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 3, spread=2.5, normals=True)
        g.glEnd()


def scene_clipping(g, W, H, seed):
    """Triangles crossing every frustum plane, including the near plane and
    vertices behind the eye."""
    rng = _rng(seed)
    _perspective(g, W, H, near=0.5, far=8.0, dist=0.0)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glBegin(E["GL_TRIANGLES"])
    for _ in range(40):
        for _v in range(3):
            g.glColor4f(*rng.random(4).astype(np.float32))
            g.glTexCoord2f(*rng.random(2).astype(np.float32))
            p = rng.uniform(-6, 6, 3)
            p[2] = rng.uniform(-12, 2)
            g.glVertex3f(*p.astype(np.float32))
    g.glEnd()


def scene_cull(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_CULL_FACE"])
    for k, (face, front) in enumerate([("GL_BACK", "GL_CCW"), ("GL_FRONT", "GL_CCW"), ("GL_BACK", "GL_CW"),
                                       ("GL_FRONT_AND_BACK", "GL_CCW"), ("GL_FRONT", "GL_CW")]):
        g.glCullFace(E[face])
        g.glFrontFace(E[front])
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 8)
        g.glEnd()


def scene_primitives(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    for mode, nv in [("GL_TRIANGLE_STRIP", 11), ("GL_TRIANGLE_FAN", 9), ("GL_QUADS", 14), ("GL_QUAD_STRIP", 13),
                     ("GL_POLYGON", 7), ("GL_TRIANGLES", 8)]:
        g.glBegin(E[mode])
        for _ in range(nv):
            g.glColor3f(*rng.random(3).astype(np.float32))
            g.glVertex3f(*((rng.random(3) - 0.5) * 3).astype(np.float32))
        g.glEnd()


def scene_arrays(g, W, H, seed):
    """Client arrays of float and unsigned-byte components, sizes 2/3/4,
    explicit strides, colour arrays of 3 components, 1-component texcoords,
    negative first, glDrawElements with 8/16/32-bit client indices."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_TEXTURE_2D"])
    _upload_texture(g, random_rgba(8, 8, seed))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR"])
    n = 30
    pos = g.keep(((rng.random((n, 4)) - 0.5) * 3).astype(np.float32))
    colb = g.keep(rng.integers(0, 256, (n, 3), dtype=np.uint8))
    colf = g.keep((rng.random((n, 4)) * 1.4 - 0.2).astype(np.float32))
    tc = g.keep(rng.random((n, 3)).astype(np.float32))
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glEnableClientState(E["GL_TEXTURE_COORD_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 16, pos)
    g.glColorPointer(3, E["GL_UNSIGNED_BYTE"], 0, colb)
    g.glTexCoordPointer(1, E["GL_FLOAT"], 12, tc)
    g.glDrawArrays(E["GL_TRIANGLES"], 0, 12)
    g.glVertexPointer(2, E["GL_FLOAT"], 16, pos)
    g.glColorPointer(4, E["GL_FLOAT"], 0, colf)
    g.glTexCoordPointer(2, E["GL_FLOAT"], 12, tc)
    g.glDrawArrays(E["GL_TRIANGLE_STRIP"], -2, 10)
    g.glVertexPointer(4, E["GL_FLOAT"], 0, pos)
    for itype, dt in [("GL_UNSIGNED_BYTE", np.uint8), ("GL_UNSIGNED_SHORT", np.uint16), ("GL_UNSIGNED_INT", np.uint32)]:
        idx = g.keep(rng.integers(0, n, 18).astype(dt))
        g.glDrawElements(E["GL_TRIANGLES"], 18, E[itype], idx)
    g.glDisableClientState(E["GL_COLOR_ARRAY"])
    g.glColor4f(0.9, 0.1, 0.5, 1.0)
    g.glDrawArrays(E["GL_TRIANGLE_FAN"], 5, 7)


def scene_vbo(g, W, H, seed, oob_index=False):
    """Interleaved buffer objects, element buffers with a byte offset and
    glBufferSubData updates.  With oob_index, one index points past the end of
    the array buffer: the reference reads outside the buffer there (undefined),
    the device and the oracle both use the get_array_element defaults, so that
    variant is compared device against oracle only."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glEnable(E["GL_COLOR_MATERIAL"])
    n = 40
    verts = np.concatenate([(rng.random((n, 3)) - 0.5) * 3, rng.normal(size=(n, 3)), rng.random((n, 4))], 1)
    verts = g.keep(verts.astype(np.float32))
    idx = g.keep(np.concatenate([rng.integers(0, n, 30), [n + 5 if oob_index else 7, 3, 4]]).astype(np.uint16))
    bufs = np.zeros(2, np.uint32)
    g.glGenBuffers(2, bufs)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(bufs[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], verts.nbytes, verts, E["GL_STATIC_DRAW"])
    patch = g.keep((rng.random(8) * 2 - 1).astype(np.float32))
    g.glBufferSubData(E["GL_ARRAY_BUFFER"], 40 * 5, patch.nbytes, patch)
    g.glBindBuffer(E["GL_ELEMENT_ARRAY_BUFFER"], int(bufs[1]))
    g.glBufferData(E["GL_ELEMENT_ARRAY_BUFFER"], idx.nbytes, idx, E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 40, 0)
    g.glNormalPointer(E["GL_FLOAT"], 40, 12)
    g.glColorPointer(4, E["GL_FLOAT"], 40, 24)
    g.glDrawElements(E["GL_TRIANGLES"], 33, E["GL_UNSIGNED_SHORT"], 0)
    g.glDrawElements(E["GL_TRIANGLES"], 15, E["GL_UNSIGNED_SHORT"], 6)
    g.glDrawArrays(E["GL_QUADS"], 2, 12)


def scene_lists(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    base = g.glGenLists(3)
    g.glNewList(base, E["GL_COMPILE"])
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 4)
    g.glEnd()
    g.glEndList()
    g.glNewList(base + 1, E["GL_COMPILE_AND_EXECUTE"])
    g.glPushMatrix()
    g.glRotatef(30.0, 0.0, 0.0, 1.0)
    g.glCallList(base)
    g.glPopMatrix()
    g.glEndList()
    g.glNewList(base + 2, E["GL_COMPILE"])
    g.glTranslatef(0.3, -0.2, 0.1)
    g.glCallList(base + 1)
    g.glEndList()
    ids = g.keep(np.array([2, 1, 2], np.uint8))
    g.glListBase(base)
    g.glCallLists(3, E["GL_UNSIGNED_BYTE"], ids)
    g.glCallList(base)


def scene_lists_replay(g, W, H, seed):
    """Display lists replayed many times: one block whose colours are all
    recorded (the same vertex records every call), one that inherits the
    current colour and normal from outside the list (records that change with
    them), a list recompiled between calls, a deleted list, lines in a list,
    and one glBegin outside a list closed by a list's glEnd (reference
    src/lists.c:185-425, src/gl_api.c:2290-2455)."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    base = g.glGenLists(5)
    g.glNewList(base, E["GL_COMPILE"])                 # fully recorded block
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 3, spread=1.5)
    g.glEnd()
    g.glEndList()
    g.glNewList(base + 1, E["GL_COMPILE"])             # inherits colour and normal
    g.glBegin(E["GL_TRIANGLES"])
    for _ in range(6):
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.5).astype(np.float32))
    g.glEnd()
    g.glEndList()
    g.glNewList(base + 2, E["GL_COMPILE"])             # lines
    g.glBegin(E["GL_LINE_STRIP"])
    for _ in range(5):
        g.glColor3f(*rng.random(3).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.5).astype(np.float32))
    g.glEnd()
    g.glEndList()
    g.glNewList(base + 3, E["GL_COMPILE"])             # vertices and the glEnd of an outer glBegin
    for _ in range(3):
        g.glColor3f(*rng.random(3).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.5).astype(np.float32))
    g.glEnd()
    g.glEndList()
    for k in range(6):
        g.glPushMatrix()
        g.glTranslatef(float(rng.uniform(-0.8, 0.8)), float(rng.uniform(-0.6, 0.6)), 0.0)
        g.glRotatef(float(rng.uniform(0, 360)), 0.0, 0.0, 1.0)
        g.glCallList(base)
        g.glColor3f(*rng.random(3).astype(np.float32))
        g.glCallList(base + 1)
        g.glCallList(base + 1)                         # same colour again
        g.glCallList(base + 2)
        g.glBegin(E["GL_TRIANGLES"])
        g.glColor3f(*rng.random(3).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.5).astype(np.float32))
        g.glCallList(base + 3)
        g.glPopMatrix()
        if k == 2:                                     # recompile: new geometry in the same list
            g.glNewList(base, E["GL_COMPILE"])
            g.glBegin(E["GL_QUADS"])
            for _ in range(8):
                g.glColor4f(*rng.random(4).astype(np.float32))
                g.glVertex3f(*((rng.random(3) - 0.5) * 1.5).astype(np.float32))
            g.glEnd()
            g.glEndList()
        if k == 4:
            g.glDeleteLists(base + 2, 1)


def _lists_repeat_blocks(g, rng):
    """The calls of one list: a triangle block that sets colour and normal
    before every vertex (inherits the texcoord), a translation, a strip whose
    first vertex inherits colour and normal, and textured quads that inherit
    the colour and set a colour after their last vertex."""
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 4, spread=1.2, normals=True)
    g.glEnd()
    g.glTranslatef(0.2, -0.1, 0.0)
    g.glBegin(E["GL_TRIANGLE_STRIP"])
    g.glTexCoord2f(0.25, 0.75)
    g.glVertex3f(-0.6, -0.5, 0.1)
    g.glColor3f(*rng.random(3).astype(np.float32))
    g.glVertex3f(-0.4, 0.5, -0.1)
    g.glNormal3f(*rng.normal(size=3).astype(np.float32))
    for _ in range(4):
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.2).astype(np.float32))
    g.glEnd()
    g.glBegin(E["GL_QUADS"])
    for _ in range(8):
        g.glTexCoord2f(*rng.random(2).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 1.2).astype(np.float32))
    g.glColor3f(*rng.random(3).astype(np.float32))
    g.glEnd()


def lists_repeat(g, W, H, seed, inline=False, probe=None):
    """A list of three blocks replayed eight times under lighting and
    GL_COLOR_MATERIAL, between changes of the current colour, normal and
    texcoord, of the colour-material mode, and a vertex left pending outside
    glBegin; after each call a triangle with no attribute calls draws with what
    the list left, every other time with GL_COLOR_MATERIAL off so that it
    shows the material the list's last vertex wrote.  inline=True issues the
    list's calls directly (what the replays must equal); probe(g) is called
    after each call of the list."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.fv(0.5, 1.0, 1.0, 0.0))
    g.glEnable(E["GL_COLOR_MATERIAL"])
    list_seed = int(rng.integers(1 << 30))
    lid = 0
    if not inline:
        lid = g.glGenLists(1)
        g.glNewList(lid, E["GL_COMPILE"])
        _lists_repeat_blocks(g, np.random.default_rng(list_seed))
        g.glEndList()
    for k in range(8):
        if k % 3 == 0:
            g.glColor3f(*rng.random(3).astype(np.float32))
        if k % 4 == 1:
            g.glNormal3f(*rng.normal(size=3).astype(np.float32))
        if k == 5:
            g.glTexCoord2f(0.5, 0.125)
        if k == 3:
            g.glColorMaterial(E["GL_FRONT"], E["GL_DIFFUSE"])
        if k == 6:
            g.glDisable(E["GL_COLOR_MATERIAL"])
        if k == 4:                                     # joins the list's first block
            g.glVertex3f(0.3, 0.3, 0.0)
        g.glPushMatrix()
        g.glTranslatef(float(rng.uniform(-0.6, 0.6)), float(rng.uniform(-0.5, 0.5)), 0.0)
        g.glRotatef(float(rng.uniform(0, 360)), 0.0, 0.0, 1.0)
        if inline:
            _lists_repeat_blocks(g, np.random.default_rng(list_seed))
        else:
            g.glCallList(lid)
        if probe is not None:
            probe(g)
        if k % 2 == 1 and k < 6:
            g.glDisable(E["GL_COLOR_MATERIAL"])
        g.glBegin(E["GL_TRIANGLES"])
        for v in ((0.0, 0.0, 0.2), (0.5, 0.1, 0.2), (0.1, 0.5, 0.2)):
            g.glVertex3f(*v)
        g.glEnd()
        if k % 2 == 1 and k < 6:
            g.glEnable(E["GL_COLOR_MATERIAL"])
        g.glPopMatrix()


def scene_lists_repeat(g, W, H, seed):
    lists_repeat(g, W, H, seed)


# screen-aligned quads (x0, row0, x1, row1, s extent, t extent) over a 256x256
# framebuffer whose LODs with a 64x64 texture span magnification to beyond the
# last level, with isotropic and anisotropic footprints
LOD_QUADS = [
    (4, 4, 67, 67, 0.5, 0.5),          # lod -0.98 (magnified)
    (72, 4, 103, 35, 0.6, 0.6),        # 0.31
    (108, 4, 139, 35, 1.2, 1.0),       # 1.31
    (144, 4, 159, 19, 1.5, 0.5),       # 2.68 (anisotropic)
    (164, 4, 171, 11, 1.3, 1.3),       # 3.57
    (176, 4, 179, 7, 3.3, 2.0),        # 6.14: the last level
    (184, 4, 185, 5, 40.0, 1.0),       # 11.3: beyond the chain
    (4, 72, 131, 103, 3.0, 0.2),       # 0.60 (s minified, t magnified)
    (136, 72, 151, 135, 0.3, 2.4),     # 1.29 (t)
    (156, 72, 219, 199, 2.7, 6.9),     # 1.80
]


def lod_quads(g, W, H, tex: np.ndarray, min_filter: str, mag_filter: str = "GL_LINEAR", quads=None):
    """Textured screen-aligned quads (GL_REPLACE, GL_REPEAT) at integer pixel
    corners under glOrtho(0, W, 0, H): the texture coordinates are affine in
    the pixel position, each quad has one LOD."""
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), -1.0, 1.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    tid = _upload_texture(g, tex)
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E[min_filter])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E[mag_filter])
    g.glEnable(E["GL_TEXTURE_2D"])
    g.glTexEnvi(E["GL_TEXTURE_ENV"], E["GL_TEXTURE_ENV_MODE"], E["GL_REPLACE"])
    g.glBegin(E["GL_QUADS"])
    for q in (LOD_QUADS if quads is None else quads):
        x0, r0, x1, r1, su, sv = q[:6]
        s0, t0 = q[6:8] if len(q) > 6 else (0.0, 0.0)   # optional texture-coordinate origin
        # GL y of a framebuffer row r is H - r
        for (x, r, s, t) in ((x0, r0, s0, t0), (x1, r0, s0 + su, t0), (x1, r1, s0 + su, t0 + sv),
                             (x0, r1, s0, t0 + sv)):
            g.glTexCoord2f(s, t)
            g.glVertex2f(float(x), float(H - r))
    g.glEnd()
    g.glDisable(E["GL_TEXTURE_2D"])
    return tid


def scene_scissor_viewport(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_SCISSOR_TEST"])
    g.glScissor(W // 5, H // 4, W // 2, H // 2)
    g.glViewport(-W // 6, H // 8, W, (3 * H) // 4)
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 10)
    g.glEnd()
    g.glClearColor(0.5, 0.25, 0.75, 1.0)
    g.glClear(E["GL_COLOR_BUFFER_BIT"])
    g.glViewport(W // 3, -H // 3, W, H * 2)
    g.glScissor(0, 0, W, H)
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 10)
    g.glEnd()


def scene_colormask(g, W, H, seed):
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_BLEND"])
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE"])
    for m in range(16):
        g.glColorMask(bool(m & 1), bool(m & 2), bool(m & 4), bool(m & 8))
        g.glBegin(E["GL_TRIANGLES"])
        _random_tris_immediate(g, rng, 2)
        g.glEnd()


def scene_edge_cases(g, W, H, seed):
    """Robustness inputs of the reference's validation demo: NaN/Inf colours,
    Inf and +-1e10 vertices, NaN texture coordinates, degenerate triangles."""
    nan, inf = float("nan"), float("inf")
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_TEXTURE_2D"])
    _upload_texture(g, random_rgba(8, 8, seed))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_LINEAR"])
    tris = [
        [(nan, 0.5, 0.2, 1.0), (-0.5, -0.5, 0.0)], [(0.2, inf, 0.3, 1.0), (0.5, -0.5, 0.0)], [(0.3, 0.4, -inf, nan), (0.0, 0.5, 0.0)],
        [(1, 0, 0, 1), (1e10, 0.0, 0.0)], [(0, 1, 0, 1), (-1e10, 1e10, 0.0)], [(0, 0, 1, 1), (0.2, 0.2, -1e10)],
        [(1, 1, 0, 1), (inf, 0.0, 0.0)], [(1, 0, 1, 1), (0.0, -inf, 0.0)], [(0, 1, 1, 1), (0.1, 0.1, 0.0)],
        [(1, 1, 1, 1), (0.0, 0.0, 0.0)], [(1, 1, 1, 1), (0.0, 0.0, 0.0)], [(1, 1, 1, 1), (0.0, 0.0, 0.0)],
        [(1, 0.5, 0, 1), (-0.8, -0.8, 0.0)], [(1, 0.5, 0, 1), (0.8, 0.8, 0.0)], [(1, 0.5, 0, 1), (0.0, 0.0, 0.0)],
    ]
    texcoords = [(nan, 0.0), (1e10, -1e10), (inf, 0.5), (0.5, nan), (-inf, 2.0)]
    g.glBegin(E["GL_TRIANGLES"])
    for k, (col, p) in enumerate(tris):
        g.glColor4f(*col)
        g.glTexCoord2f(*texcoords[k % len(texcoords)])
        g.glVertex3f(*p)
    g.glEnd()


def scene_midprimitive_state(g, W, H, seed):
    """State changed between the vertices of one primitive: matrices,
    materials, lights and the colour-material side effect (reference
    emit_vertex transforms each vertex with the state current at its call)."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glEnable(E["GL_LIGHT1"])
    g.glLightfv(E["GL_LIGHT1"], E["GL_DIFFUSE"], g.fv(0.9, 0.2, 0.1, 1.0))
    g.glLightfv(E["GL_LIGHT1"], E["GL_POSITION"], g.fv(-1.0, 2.0, 1.0, 1.0))
    g.glShadeModel(E["GL_PHONG"])
    g.glLightModeli(E["GL_LIGHT_MODEL_TWO_SIDE"], 1)
    g.glBegin(E["GL_TRIANGLES"])
    for k in range(24):
        if k % 4 == 1:
            g.glMaterialfv(E["GL_FRONT"], E["GL_DIFFUSE"], g.fv(*rng.random(4)))
        if k % 5 == 2:
            g.glRotatef(float(rng.uniform(-40, 40)), 0.0, 1.0, 0.0)
        if k % 7 == 3:
            g.glEnable(E["GL_COLOR_MATERIAL"])
        if k % 7 == 6:
            g.glDisable(E["GL_COLOR_MATERIAL"])
        if k == 11:
            g.glShadeModel(E["GL_SMOOTH"])
        g.glColor4f(*rng.random(4).astype(np.float32))
        g.glNormal3f(*rng.normal(size=3).astype(np.float32))
        g.glVertex3f(*((rng.random(3) - 0.5) * 3).astype(np.float32))
    g.glEnd()


def _prim_state(g, rng, W, H, k):
    """Per-batch fragment state shared by the line and point scenes: depth
    function and writes, blending, colour mask, scissor, fog mode.  The alpha
    test stays off: the reference's line loop never ends on a pixel it
    rejects."""
    funcs = ["GL_LESS", "GL_LEQUAL", "GL_GREATER", "GL_ALWAYS", "GL_NOTEQUAL"]
    g.glDepthFunc(E[funcs[k % len(funcs)]])
    g.glDepthMask(int(rng.random() < 0.7))
    (g.glEnable if rng.random() < 0.4 else g.glDisable)(E["GL_BLEND"])
    g.glBlendFunc(E[["GL_SRC_ALPHA", "GL_ONE", "GL_DST_COLOR", "GL_ONE_MINUS_SRC_ALPHA"][rng.integers(4)]],
                  E[["GL_ONE_MINUS_SRC_ALPHA", "GL_ZERO", "GL_ONE", "GL_SRC_COLOR"][rng.integers(4)]])
    if rng.random() < 0.25:
        g.glColorMask(*(int(v) for v in rng.random(4) < 0.6))
    else:
        g.glColorMask(1, 1, 1, 1)
    if rng.random() < 0.3:
        g.glEnable(E["GL_SCISSOR_TEST"])
        g.glScissor(int(rng.integers(0, max(1, W // 2))), int(rng.integers(0, max(1, H // 2))),
                    int(rng.integers(W // 4, max(W // 4 + 1, W))), int(rng.integers(H // 4, max(H // 4 + 1, H))))
    else:
        g.glDisable(E["GL_SCISSOR_TEST"])
    if rng.random() < 0.4:
        g.glEnable(E["GL_FOG"])
        g.glFogi(E["GL_FOG_MODE"], E[["GL_LINEAR", "GL_EXP", "GL_EXP2"][rng.integers(3)]])
        g.glFogf(E["GL_FOG_DENSITY"], float(rng.uniform(0.05, 0.6)))
        g.glFogf(E["GL_FOG_START"], -5.0)           # lines and points fog on the negated eye_z
        g.glFogf(E["GL_FOG_END"], -1.0)
    else:
        g.glDisable(E["GL_FOG"])


def scene_lines_points(g, W, H, seed):
    """GL_POINTS (sizes 1 to 5, textured and alpha-tested), GL_LINES,
    GL_LINE_STRIP and GL_LINE_LOOP (widths 1 to 4.5, textured, crossing the
    near plane and the sides of the frustum), each batch under its own depth,
    blend, colour-mask, scissor and fog state (reference src/raster.c:65-296,
    :1019-1196)."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 4)
    g.glEnd()
    tex = _upload_texture(g, random_rgba(16, 16, seed + 90))
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"],
                      E["GL_LINEAR"] if seed % 2 else E["GL_NEAREST"])
    g.glBindTexture(E["GL_TEXTURE_2D"], tex)
    modes = ["GL_POINTS", "GL_LINES", "GL_LINE_STRIP", "GL_LINE_LOOP"]
    for k in range(16):
        mode = modes[k % 4]
        _prim_state(g, rng, W, H, k)
        (g.glEnable if rng.random() < 0.4 else g.glDisable)(E["GL_TEXTURE_2D"])
        if mode == "GL_POINTS":
            g.glPointSize(float(rng.choice([1.0, 2.4, 3.6, 5.0])))
            if rng.random() < 0.5:
                g.glEnable(E["GL_ALPHA_TEST"])
                g.glAlphaFunc(E[["GL_GREATER", "GL_LESS", "GL_GEQUAL"][rng.integers(3)]], float(rng.random()))
        else:
            g.glLineWidth(float(rng.choice([1.0, 2.0, 3.0, 4.5])))
        spread = 5.0 if k % 3 == 0 else 2.5          # some primitives leave the frustum
        g.glBegin(E[mode])
        for _ in range(int(rng.integers(2, 9))):
            g.glColor4f(*rng.random(4).astype(np.float32))
            g.glTexCoord2f(*rng.uniform(-1.0, 2.0, 2).astype(np.float32))
            v = (rng.random(3) - 0.5) * spread
            if k % 4 == 1:
                v[2] = rng.uniform(1.5, 3.5)             # through the near plane (eye at z = 3)
            g.glVertex3f(*v.astype(np.float32))
        g.glEnd()
        g.glDisable(E["GL_ALPHA_TEST"])
    g.glColorMask(1, 1, 1, 1)
    g.glDepthMask(1)


def setup_lines_points(g, W, H, seed=SEED_BASE + 9, n_lines=20_000, n_points=20_000):
    """State and one interleaved buffer object of the many-lines-and-points
    workload: n_lines short segments (first endpoint and n_points points
    uniform over a box that overhangs the frustum, so some segments are
    clipped; second endpoint N(0, 0.04) away, tens of pixels at 1080p) and
    random colours."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    nv = 2 * n_lines + n_points
    v = np.empty((nv, 8), np.float32)                    # x y z, r g b a, pad
    v[:, 0:3] = ((rng.random((nv, 3)) - 0.5) * np.array([3.2, 2.0, 2.0])).astype(np.float32)
    v[1:2 * n_lines:2, 0:3] = v[0:2 * n_lines:2, 0:3] + rng.normal(0.0, 0.04, (n_lines, 3)).astype(np.float32)
    v[:, 3:7] = rng.random((nv, 4)).astype(np.float32)
    v[:, 7] = 0.0
    ids = np.zeros(1, np.uint32)
    g.glGenBuffers(1, ids)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(ids[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], v.nbytes, g.keep(v), E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_COLOR_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 32, 0)
    g.glColorPointer(4, E["GL_FLOAT"], 32, 12)
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
    return n_lines, n_points


def draw_lines_points(g, st):
    """GL_LINES of width 2, then GL_POINTS of size 3 blended, under the depth
    test."""
    n_lines, n_points = st
    g.glLineWidth(2.0)
    g.glDrawArrays(E["GL_LINES"], 0, 2 * n_lines)
    g.glEnable(E["GL_BLEND"])
    g.glPointSize(3.0)
    g.glDrawArrays(E["GL_POINTS"], 2 * n_lines, n_points)
    g.glDisable(E["GL_BLEND"])


def lines_points_workload(g, W, H, seed=SEED_BASE + 9, n_lines=20_000, n_points=20_000):
    """setup_lines_points and one draw_lines_points."""
    st = setup_lines_points(g, W, H, seed, n_lines, n_points)
    draw_lines_points(g, st)
    g.glDisableClientState(E["GL_VERTEX_ARRAY"])
    g.glDisableClientState(E["GL_COLOR_ARRAY"])
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], 0)
    return 2 * st[0] + st[1]


def scene_prim_edges(g, W, H, seed):
    """Edge cases of points and lines: zero-length segments, a GL_LINES
    block with an odd vertex count, a one-vertex GL_LINE_LOOP and
    GL_LINE_STRIP, widths and sizes below one (read as one) and large ones,
    segments leaving every side of the frustum, points on the frustum's
    edges and behind the eye, lines and points under a 1:1 orthographic
    projection at integer and half-integer positions, and a depth range
    other than [0, 1] (reference src/raster.c:107-296, :1019-1196)."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glDepthRange(0.2, 0.7)

    def vtx(v, c=None):
        g.glColor4f(*(rng.random(4) if c is None else c).astype(np.float32))
        g.glVertex3f(*np.asarray(v, np.float32))
    for width in (0.3, 1.0, 6.0):
        g.glLineWidth(width)
        g.glBegin(E["GL_LINES"])
        vtx((0.1, 0.1, 0.0)); vtx((0.1, 0.1, 0.0))                 # zero length
        vtx((-5.0, 0.2, 0.0)); vtx((5.0, -0.3, 0.5))               # left and right planes
        vtx((0.3, -4.0, 0.2)); vtx((-0.2, 4.0, -0.4))              # bottom and top
        vtx((0.0, 0.0, 2.95)); vtx((0.4, 0.3, -50.0))              # near and far
        vtx((0.5, 0.5, 0.0))                                       # odd vertex: ignored
        g.glEnd()
    for mode in ("GL_LINE_LOOP", "GL_LINE_STRIP"):
        g.glBegin(E[mode])
        vtx((0.2, -0.2, 0.0))                                      # one vertex: nothing
        g.glEnd()
    for size in (0.2, 1.0, 9.0):
        g.glPointSize(size)
        g.glBegin(E["GL_POINTS"])
        for _ in range(6):
            vtx((rng.random(3) - 0.5) * 2.4)
        vtx((0.0, 0.0, 3.5))                                       # behind the eye
        vtx((1.0e-7, 0.0, 2.9))                                    # just in front of the near plane
        g.glEnd()
    g.glDepthRange(0.0, 1.0)
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), -1.0, 1.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    g.glLineWidth(1.0)
    g.glBegin(E["GL_LINE_LOOP"])
    for x, y in ((2.0, 2.0), (W - 3.0, 2.5), (W - 3.5, H - 3.0), (2.5, H - 2.5)):
        vtx((x, y, 0.0))
    g.glEnd()
    g.glPointSize(2.0)
    g.glBegin(E["GL_POINTS"])
    for x, y in ((0.0, 0.0), (W - 1.0, H - 1.0), (W / 2 + 0.5, H / 2 + 0.5), (-1.0, 5.0)):
        vtx((x, y, 0.0))
    g.glEnd()


def _polygon_mode_tris(g, rng, W, H, seed, modes):
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glEnable(E["GL_COLOR_MATERIAL"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.keep(np.array([1.0, 1.0, 1.0, 0.0], np.float32)))
    g.glPointSize(3.0)                                # draw_triangle_points ignores it
    g.glLineWidth(2.0)
    for k, (front, back) in enumerate(modes):
        g.glPolygonMode(E["GL_FRONT"], E[front])
        g.glPolygonMode(E["GL_BACK"], E[back])
        g.glShadeModel(E["GL_PHONG"] if k % 2 else E["GL_SMOOTH"])
        (g.glEnable if k % 3 == 2 else g.glDisable)(E["GL_CULL_FACE"])
        _prim_state(g, rng, W, H, k)
        if front == "GL_POINT" and back in ("GL_POINT", "GL_FILL") or back == "GL_POINT" and front == "GL_FILL":
            g.glEnable(E["GL_ALPHA_TEST"])               # never on lines: see _prim_state
            g.glAlphaFunc(E["GL_GREATER"], 0.3)
        g.glBegin(E[["GL_TRIANGLES", "GL_QUADS", "GL_TRIANGLE_FAN"][k % 3]])
        _random_tris_immediate(g, rng, 4, spread=4.0, normals=True)
        g.glEnd()
        g.glDisable(E["GL_ALPHA_TEST"])
    g.glColorMask(1, 1, 1, 1)
    g.glDepthMask(1)


def scene_polygon_mode(g, W, H, seed):
    """glPolygonMode LINE and POINT on both faces (GL_PHONG corners lit per
    vertex with the front material, culling before the mode choice;
    reference src/raster.c:841-898, :934-952)."""
    modes = [("GL_LINE", "GL_LINE"), ("GL_POINT", "GL_POINT"), ("GL_LINE", "GL_POINT"), ("GL_POINT", "GL_LINE"),
             ("GL_LINE", "GL_LINE"), ("GL_POINT", "GL_POINT")]
    _polygon_mode_tris(g, _rng(seed), W, H, seed, modes)


def scene_polygon_mode_mixed(g, W, H, seed):
    """glPolygonMode with one face filled and the other drawn as lines or
    points in the same draw (reference src/raster.c:936-957): the mode is
    chosen per fan triangle, so filled, wireframe and point fragments of one
    draw meet on the same pixels in submission order."""
    modes = [("GL_FILL", "GL_LINE"), ("GL_POINT", "GL_FILL"), ("GL_LINE", "GL_FILL"), ("GL_FILL", "GL_POINT")]
    _polygon_mode_tris(g, _rng(seed), W, H, seed, modes)


def scene_polygon_mode_mixed_dense(g, W, H, seed):
    """Many overlapping triangles per mixed-mode draw, so that filled faces
    and the other faces' lines and points cross on most pixels: a stencil
    test passing up to three layers (GL_GEQUAL 2, GL_INCR) under blending, a textured draw whose filled faces
    take the alpha test on the texel (points too; lines never, as the
    reference's line loop does not end on a rejected pixel), GL_PHONG with
    fog, and culling of one face."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glClearStencil(0)
    g.glClear(E["GL_STENCIL_BUFFER_BIT"])
    tex = random_rgba(16, 16, seed + 5)
    tex[..., 3] = rng.integers(0, 256, (16, 16), dtype=np.uint8)
    _upload_texture(g, tex)
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR_MIPMAP_LINEAR"])
    g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_LINEAR"])
    g.glLineWidth(2.0)
    g.glPointSize(2.0)
    # 1: front filled, back wireframe, stencil GL_GEQUAL 2 / GL_INCR (up to
    # three layers), blending
    g.glEnable(E["GL_STENCIL_TEST"])
    g.glStencilFunc(E["GL_GEQUAL"], 2, 0xFF)
    g.glStencilOp(E["GL_KEEP"], E["GL_INCR"], E["GL_INCR"])
    g.glEnable(E["GL_BLEND"])
    g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
    g.glPolygonMode(E["GL_FRONT"], E["GL_FILL"])
    g.glPolygonMode(E["GL_BACK"], E["GL_LINE"])
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 60, spread=2.0, alpha=(0.3, 1.0))
    g.glEnd()
    g.glDisable(E["GL_BLEND"])
    g.glDisable(E["GL_STENCIL_TEST"])
    # 2: textured; front points, back filled; the alpha test on the texel
    g.glEnable(E["GL_TEXTURE_2D"])
    g.glEnable(E["GL_ALPHA_TEST"])
    g.glAlphaFunc(E["GL_GREATER"], 0.4)
    g.glPolygonMode(E["GL_FRONT"], E["GL_POINT"])
    g.glPolygonMode(E["GL_BACK"], E["GL_FILL"])
    g.glBegin(E["GL_QUADS"])
    _random_tris_immediate(g, rng, 40, spread=2.0, tex=True)
    g.glEnd()
    g.glDisable(E["GL_ALPHA_TEST"])
    # 3: textured, front filled, back wireframe, GL_LEQUAL, linear fog
    g.glDepthFunc(E["GL_LEQUAL"])
    g.glEnable(E["GL_FOG"])
    g.glFogi(E["GL_FOG_MODE"], E["GL_LINEAR"])
    g.glFogf(E["GL_FOG_START"], 1.0)
    g.glFogf(E["GL_FOG_END"], 6.0)
    g.glPolygonMode(E["GL_FRONT"], E["GL_FILL"])
    g.glPolygonMode(E["GL_BACK"], E["GL_LINE"])
    g.glBegin(E["GL_TRIANGLE_STRIP"])
    _random_tris_immediate(g, rng, 30, spread=2.5, tex=True)
    g.glEnd()
    g.glDisable(E["GL_FOG"])
    g.glDisable(E["GL_TEXTURE_2D"])
    # 4: GL_PHONG, front wireframe, back filled; back faces culled, then front
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glEnable(E["GL_COLOR_MATERIAL"])
    g.glShadeModel(E["GL_PHONG"])
    g.glPolygonMode(E["GL_FRONT"], E["GL_LINE"])
    g.glPolygonMode(E["GL_BACK"], E["GL_FILL"])
    for face in ("GL_BACK", "GL_FRONT"):
        g.glEnable(E["GL_CULL_FACE"])
        g.glCullFace(E[face])
        g.glBegin(E["GL_TRIANGLE_FAN"])
        _random_tris_immediate(g, rng, 12, spread=2.5, normals=True)
        g.glEnd()
    g.glDisable(E["GL_CULL_FACE"])
    g.glDisable(E["GL_LIGHTING"])
    g.glShadeModel(E["GL_SMOOTH"])
    g.glPolygonMode(E["GL_FRONT_AND_BACK"], E["GL_FILL"])
    g.glDepthFunc(E["GL_LESS"])


def scene_heavy_bins(g, W, H, seed):
    """Fans of 300 thin triangles around a few points, so that the tiles
    there hold bins of several hundred entries (the device rasterises such
    tiles one pixel row at a time), under every depth function with and
    without depth writes, with the stencil test, and textured."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    funcs = ["GL_LESS", "GL_LEQUAL", "GL_GREATER", "GL_GEQUAL", "GL_EQUAL", "GL_NOTEQUAL", "GL_ALWAYS", "GL_NEVER"]
    for k, fn in enumerate(funcs + ["GL_LESS", "GL_GREATER"]):
        g.glDepthFunc(E[fn])
        g.glDepthMask(int(k % 3 != 1))
        if k == 8:          # stencil on: the rows keep the in-order loop
            g.glEnable(E["GL_STENCIL_TEST"])
            g.glStencilFunc(E["GL_GEQUAL"], 3, 0xFF)
            g.glStencilOp(E["GL_INCR"], E["GL_KEEP"], E["GL_INCR"])
        if k == 9:          # textured, stencil off again
            g.glDisable(E["GL_STENCIL_TEST"])
            g.glEnable(E["GL_TEXTURE_2D"])
            _upload_texture(g, random_rgba(8, 8, seed + 11))
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        g.glBegin(E["GL_TRIANGLES"])
        for t in range(300):
            a = rng.uniform(0.0, 2.0 * np.pi)
            da = rng.uniform(0.02, 0.3)
            r = rng.uniform(0.2, 1.6)
            z = float(rng.uniform(-1.0, 1.0)) if k % 4 else float(rng.choice([-0.5, 0.0, 0.5]))
            for (px, py, pz) in ((cx, cy, z), (cx + r * np.cos(a), cy + r * np.sin(a), z + 0.3),
                                 (cx + r * np.cos(a + da), cy + r * np.sin(a + da), z - 0.3)):
                g.glColor4f(*rng.random(4).astype(np.float32))
                g.glTexCoord2f(*rng.uniform(0.0, 1.0, 2).astype(np.float32))
                g.glVertex3f(float(px), float(py), float(pz))
        g.glEnd()
    g.glDisable(E["GL_TEXTURE_2D"])
    g.glDepthMask(1)
    g.glDepthFunc(E["GL_LESS"])


def scene_draw_pixels(g, W, H, seed):
    """glDrawPixels over rendered triangles (reference src/gl_api.c:1284-1373):
    all four formats, rectangles hanging off every edge, alpha test, depth
    test against the triangles (incoming depth 0) with and without depth
    writes, every blend factor, and an invalid raster position (behind the
    eye), which draws nothing."""
    rng = _rng(seed)
    _perspective(g, W, H)
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glBegin(E["GL_TRIANGLES"])
    _random_tris_immediate(g, rng, 6)
    g.glEnd()
    g.glRasterPos3f(0.0, 0.0, 5.0)        # eye z = +2: clip w <= 0, raster position invalid
    g.glDrawPixels(4, 4, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], g.keep(np.full(64, 255, np.uint8)))
    g.glMatrixMode(E["GL_PROJECTION"])
    g.glLoadIdentity()
    g.glOrtho(0.0, float(W), 0.0, float(H), -1.0, 1.0)
    g.glMatrixMode(E["GL_MODELVIEW"])
    g.glLoadIdentity()
    fmts = [("GL_RGBA", 4), ("GL_RGB", 3), ("GL_LUMINANCE", 1), ("GL_LUMINANCE_ALPHA", 2)]
    funcs = ["GL_NEVER", "GL_LESS", "GL_EQUAL", "GL_LEQUAL", "GL_GREATER", "GL_NOTEQUAL", "GL_GEQUAL", "GL_ALWAYS"]
    factors = ["GL_ZERO", "GL_ONE", "GL_SRC_COLOR", "GL_ONE_MINUS_SRC_COLOR", "GL_DST_COLOR",
               "GL_ONE_MINUS_DST_COLOR", "GL_SRC_ALPHA", "GL_ONE_MINUS_SRC_ALPHA", "GL_DST_ALPHA",
               "GL_ONE_MINUS_DST_ALPHA", "GL_SRC_ALPHA_SATURATE", "GL_CONSTANT_COLOR"]
    for k in range(20):
        fmt, bpp = fmts[k % 4]
        w, h = int(rng.integers(1, max(2, W // 2 + 1))), int(rng.integers(1, max(2, H // 2 + 1)))
        x, y = int(rng.integers(-w // 2, W - w // 2)), int(rng.integers(-h // 2, H - h // 2))
        if k % 5 == 4:
            g.glRasterPos2f(x + 0.37, y + 0.81)
        else:
            g.glRasterPos2i(x, y)
        (g.glEnable if rng.random() < 0.5 else g.glDisable)(E["GL_ALPHA_TEST"])
        g.glAlphaFunc(E[funcs[rng.integers(8)]], float(rng.choice([0.0, 0.5, 1.0, rng.random()])))
        (g.glEnable if rng.random() < 0.6 else g.glDisable)(E["GL_DEPTH_TEST"])
        g.glDepthFunc(E[funcs[rng.integers(8)]])
        g.glDepthMask(int(rng.random() < 0.5))
        (g.glEnable if rng.random() < 0.6 else g.glDisable)(E["GL_BLEND"])
        g.glBlendFunc(E[factors[rng.integers(len(factors))]],
                      E[[f for f in factors if f != "GL_SRC_ALPHA_SATURATE"][rng.integers(len(factors) - 1)]])
        data = rng.integers(0, 256, size=w * h * bpp, dtype=np.uint8)
        if k % 7 == 3:
            data[::bpp] = 255                 # a luminance / red channel at full scale
        g.glDrawPixels(w, h, E[fmt], E["GL_UNSIGNED_BYTE"], g.keep(data))
    g.glDepthMask(1)


def scene_vbo_oob(g, W, H, seed):
    scene_vbo(g, W, H, seed, oob_index=True)


def scene_icosphere_small(g, W, H, seed):
    draw_icosphere(g, setup_icosphere(g, W, H, seed))


def scene_textured_sphere_small(g, W, H, seed):
    draw_textured_sphere(g, setup_textured_sphere(g, W, H, seed, grid=24, tex_size=64))


def scene_overdraw_small(g, W, H, seed):
    draw_overdraw(g, setup_overdraw(g, W, H, seed, ntri=3000, max_log_area=3.0))


# ---------------------------------------------------------------------------
# batched workload: many small independent scenes, one context each
# ---------------------------------------------------------------------------

def batch_specs(n: int = 512, seed: int = SEED_BASE + 5, fb=(64, 1024), tris=(100, 50_000), tex=(64, 512)):
    """n scene descriptions: square framebuffer side log-uniform in fb; a
    lat-long grid sphere of 2 g^2 triangles with 2 g^2 log-uniform in tris;
    every other scene textured with a power-of-two texture whose side is
    log-uniform in tex; random yaw and distance."""
    rng = _rng(seed)
    lt = int(round(math.log2(tex[0]))), int(round(math.log2(tex[1])))
    out = []
    for i in range(n):
        side = int(round(math.exp(rng.uniform(math.log(fb[0]), math.log(fb[1])))))
        t = math.exp(rng.uniform(math.log(tris[0]), math.log(tris[1])))
        grid = max(4, int(round(math.sqrt(t / 2.0))))
        out.append({
            "w": side, "h": side, "grid": grid, "triangles": 2 * grid * grid,
            "textured": i % 2 == 0, "tex": 2 ** int(rng.integers(lt[0], lt[1] + 1)),
            "seed": int(seed * 1000 + i), "yaw": float(rng.uniform(0.0, 360.0)),
            "dist": float(rng.uniform(2.2, 3.5)), "clear": [float(c) for c in rng.random(3)],
        })
    return out


_TEX_CACHE = {}


def _batch_texture(size: int) -> np.ndarray:
    if size not in _TEX_CACHE:
        _TEX_CACHE[size] = np.ascontiguousarray(value_noise_rgba(size, SEED_BASE + 6 + size))
    return _TEX_CACHE[size]


def setup_batch_scene(g, spec):
    """State, texture and buffer objects of one batched scene on the current
    context; returns the index count of its draw."""
    W, H = spec["w"], spec["h"]
    verts, idx = sphere_grid(spec["grid"], spec["seed"])
    g.glViewport(0, 0, W, H)
    g.glClearColor(spec["clear"][0], spec["clear"][1], spec["clear"][2], 1.0)
    _perspective(g, W, H, dist=-spec["dist"])
    g.glRotatef(spec["yaw"], 0.0, 1.0, 0.0)
    g.glEnable(E["GL_LIGHTING"])
    g.glEnable(E["GL_LIGHT0"])
    g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.fv(1, 1, 1, 0))
    g.glEnable(E["GL_DEPTH_TEST"])
    g.glHint(E["GL_PERSPECTIVE_CORRECTION_HINT"], E["GL_NICEST"])
    if spec["textured"]:
        tex = _batch_texture(spec["tex"])
        ids = np.zeros(1, np.uint32)
        g.glGenTextures(1, ids)
        g.glBindTexture(E["GL_TEXTURE_2D"], int(ids[0]))
        g.glTexImage2D(E["GL_TEXTURE_2D"], 0, E["GL_RGBA"], spec["tex"], spec["tex"], 0, E["GL_RGBA"],
                       E["GL_UNSIGNED_BYTE"], g.keep(tex))
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MIN_FILTER"], E["GL_LINEAR_MIPMAP_LINEAR"])
        g.glTexParameteri(E["GL_TEXTURE_2D"], E["GL_TEXTURE_MAG_FILTER"], E["GL_LINEAR"])
        g.glEnable(E["GL_TEXTURE_2D"])
    else:
        g.glEnable(E["GL_COLOR_MATERIAL"])
        g.glColor4f(0.8, 0.6, 0.3, 1.0)
    bufs = np.zeros(2, np.uint32)
    g.glGenBuffers(2, bufs)
    g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(bufs[0]))
    g.glBufferData(E["GL_ARRAY_BUFFER"], verts.nbytes, g.keep(verts), E["GL_STATIC_DRAW"])
    g.glBindBuffer(E["GL_ELEMENT_ARRAY_BUFFER"], int(bufs[1]))
    g.glBufferData(E["GL_ELEMENT_ARRAY_BUFFER"], idx.nbytes, g.keep(idx), E["GL_STATIC_DRAW"])
    g.glEnableClientState(E["GL_VERTEX_ARRAY"])
    g.glEnableClientState(E["GL_NORMAL_ARRAY"])
    g.glVertexPointer(3, E["GL_FLOAT"], 32, 0)
    g.glNormalPointer(E["GL_FLOAT"], 32, 12)
    if spec["textured"]:
        g.glEnableClientState(E["GL_TEXTURE_COORD_ARRAY"])
        g.glTexCoordPointer(2, E["GL_FLOAT"], 32, 24)
    return len(idx)


def frame_batch_scene(g, n):
    """One frame of a batched scene: clear colour and depth, draw the sphere."""
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
    g.glDrawElements(E["GL_TRIANGLES"], n, E["GL_UNSIGNED_INT"], 0)


class BatchRun:
    """Contexts of a batched workload on one GLLibrary.  items are
    (W, H, prepare, frame): prepare(g) runs at once on the new context and
    returns what frame(g, state) needs; frame issues one frame's GL calls."""

    def __init__(self, g, items):
        self.g = g
        self.items = items
        self.ctxs, self.states = [], []
        for W, H, prepare, _ in items:
            ctx = g.create_context(W, H)
            self.ctxs.append(ctx)
            self.states.append(prepare(g) if prepare else None)
        self.batch = None

    def frames_immediate(self):
        """One frame of every context through the ordinary per-draw path."""
        for ctx, (_, _, _, frame), st in zip(self.ctxs, self.items, self.states):
            self.g.gl_make_current(ctx)
            frame(self.g, st)

    def record(self, autotune=True):
        """Record one frame of every context into a new batch and build it."""
        b = self.g.batch()
        b.begin()
        try:
            self.frames_immediate()
        finally:
            b.end(autotune)
        self.batch = b
        return b

    def finish(self):
        for ctx in self.ctxs:
            self.g.gl_make_current(ctx)
            self.g.glFinish()

    def read(self, i):
        W, H = self.items[i][0], self.items[i][1]
        return self.g.read_framebuffer(self.ctxs[i], W, H)

    def close(self):
        if self.batch is not None:
            self.batch.destroy()
            self.batch = None
        for ctx in self.ctxs:
            self.g.destroy_context(ctx)
        self.ctxs = []
        self.g.release()


def batch_items(specs):
    """BatchRun items of batch_specs() scenes."""
    return [(sp["w"], sp["h"], (lambda g, sp=sp: setup_batch_scene(g, sp)), frame_batch_scene) for sp in specs]


def feature_items(names, seed, size=(96, 72)):
    """BatchRun items of feature scenes: each frame clears every plane, then
    runs the scene (state calls take effect at once, clears and draws are
    what a batch records)."""
    def frame_of(fn, W, H):
        def frame(g, _):
            g.glClearColor(0.0, 0.0, 0.0, 1.0)
            g.glClearDepth(1.0)
            g.glClearStencil(0)
            g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
            fn(g, W, H, seed)
        return frame
    out = []
    for name in names:
        W, H = (160, 120) if "small" in name else size
        out.append((W, H, None, frame_of(FEATURE_SCENES[name], W, H)))
    return out


# scenes whose reference behaviour is undefined (reads outside a buffer); they
# are compared device against oracle only
UNDEFINED_IN_REFERENCE = ("vbo_oob",)
# scenes with calls a batch does not record (glDrawPixels, points, lines and
# polygon-mode LINE / POINT draws are refused while recording) or whose draws
# read display-list segments, which a batch does not track; they stay out of
# the batched feature tests
NOT_BATCHABLE = ("draw_pixels", "lines_points", "polygon_mode", "polygon_mode_mixed", "polygon_mode_mixed_dense",
                 "lists_replay", "lists_repeat", "prim_edges")
# scenes with draws the device refuses; oracle against reference only
DEVICE_UNSUPPORTED = ()

FEATURE_SCENES = {
    "depth_funcs": scene_depth_funcs,
    "blend": scene_blend,
    "stencil": scene_stencil,
    "textures": scene_textures,
    "fog": scene_fog,
    "lighting": scene_lighting,
    "clipping": scene_clipping,
    "cull": scene_cull,
    "primitives": scene_primitives,
    "arrays": scene_arrays,
    "vbo": scene_vbo,
    "vbo_oob": scene_vbo_oob,
    "draw_pixels": scene_draw_pixels,
    "lines_points": scene_lines_points,
    "prim_edges": scene_prim_edges,
    "polygon_mode": scene_polygon_mode,
    "polygon_mode_mixed": scene_polygon_mode_mixed,
    "polygon_mode_mixed_dense": scene_polygon_mode_mixed_dense,
    "heavy_bins": scene_heavy_bins,
    "lists": scene_lists,
    "lists_replay": scene_lists_replay,
    "lists_repeat": scene_lists_repeat,
    "scissor_viewport": scene_scissor_viewport,
    "colormask": scene_colormask,
    "edge_cases": scene_edge_cases,
    "midprimitive_state": scene_midprimitive_state,
    "icosphere_small": scene_icosphere_small,
    "textured_sphere_small": scene_textured_sphere_small,
    "overdraw_small": scene_overdraw_small,
}
