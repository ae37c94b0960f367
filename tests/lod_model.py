"""An independent model of the MTGL_LOD_FULL_CHAIN extension (full mip chain,
per-pixel LOD; include/mtgl_dev.h) for screen-aligned textured quads under an
orthographic projection, where the texture coordinates are affine in the
pixel position and the LOD of a quad is a closed form:
lod = log2(max(texels per pixel along x, along y)).

The chain is built as the box filter of each level (((c00 + c10) + c01) + c11)
* 0.25 on bytes / 255 in fp32, packed with truncation (the reference's level
1, src/textures.c:311-354, applied again per level); sampling follows the
reference's level sampling (src/textures.c:362-456: bilinear rows then
columns, RGBA8 truncation; nearest with a clamped index) in fp32, and level
selection follows GL 1.5 (nearest level ceil(lod + 0.5) - 1; linear between
floor(lod) and the next, blended like the reference's level 0 / level 1
blend).  Texture coordinates are exact (fp64, rounded once to fp32) where the
device interpolates them in fp32, so the two can differ by an RGBA8 step
where a bilinear tap sits on a texel boundary, and by a whole texel where a
nearest tap does: render() marks the latter pixels (AMBIGUOUS_TEXELS wide).
"""
from __future__ import annotations

import numpy as np

F = np.float32
AMBIGUOUS_TEXELS = 1e-3     # a nearest tap this close to a texel edge may round either way


def pack(x):
    x = np.clip(np.nan_to_num(x.astype(F), nan=0.0), F(0), F(1))
    return (x * F(255)).astype(np.uint8)


def unpack(b):
    return b.astype(F) / F(255)


def box_chain(level0: np.ndarray):
    """Levels of the chain: (h, w, 4) uint8 arrays, level 0 first, while both
    sides of the previous level are >= 2."""
    levels = [level0]
    cur = level0
    while cur.shape[0] >= 2 and cur.shape[1] >= 2:
        h2, w2 = cur.shape[0] // 2, cur.shape[1] // 2
        c = unpack(cur[:2 * h2, :2 * w2])
        s = ((c[0::2, 0::2] + c[0::2, 1::2]) + c[1::2, 0::2]) + c[1::2, 1::2]
        cur = pack(s * F(0.25))
        levels.append(cur)
    return levels


def wrap_uv(u, v, repeat=True):
    if repeat:
        u = u - np.trunc(u).astype(F)
        u = np.where(u < 0, u + F(1), u)
        v = v - np.trunc(v).astype(F)
        v = np.where(v < 0, v + F(1), v)
    else:
        u, v = np.clip(u, F(0), F(1)), np.clip(v, F(0), F(1))
    return u.astype(F), v.astype(F)


def nearest_ambiguous(L: np.ndarray, u, v):
    h, w = L.shape[:2]
    a, b = u.astype(np.float64) * w, v.astype(np.float64) * h
    return (np.abs(a - np.round(a)) < AMBIGUOUS_TEXELS) | (np.abs(b - np.round(b)) < AMBIGUOUS_TEXELS)


def sample_level(L: np.ndarray, u, v, linear: bool):
    h, w = L.shape[:2]
    tx = u * F(w) - F(0.5)
    ty = v * F(h) - F(0.5)
    if linear:
        x0 = np.floor(tx).astype(np.int64)
        y0 = np.floor(ty).astype(np.int64)
        fx = (tx - x0.astype(F))[..., None]
        fy = (ty - y0.astype(F))[..., None]
        xa, xb = x0 % w, (x0 + 1) % w
        ya, yb = y0 % h, (y0 + 1) % h
        c00, c10, c01, c11 = unpack(L[ya, xa]), unpack(L[ya, xb]), unpack(L[yb, xa]), unpack(L[yb, xb])
        top = c00 * (F(1) - fx) + c10 * fx
        bot = c01 * (F(1) - fx) + c11 * fx
        return pack(top * (F(1) - fy) + bot * fy)
    x = np.clip(np.floor(tx + F(0.5)).astype(np.int64), 0, w - 1)
    y = np.clip(np.floor(ty + F(0.5)).astype(np.int64), 0, h - 1)
    return L[y, x]


def sample_chain(levels, u, v, lod: float, min_filter: str, mag_filter: str):
    """(packed texels, pixels whose nearest tap is ambiguous)"""
    u, v = wrap_uv(u, v)
    none = np.zeros(u.shape, bool)

    def one(L, lin):
        return sample_level(L, u, v, lin), (none if lin else nearest_ambiguous(L, u, v))
    if not lod > 0.0:
        return one(levels[0], mag_filter == "GL_LINEAR")
    if min_filter in ("GL_NEAREST", "GL_LINEAR"):
        return one(levels[0], min_filter == "GL_LINEAR")
    lin = min_filter in ("GL_LINEAR_MIPMAP_NEAREST", "GL_LINEAR_MIPMAP_LINEAR")
    top = len(levels) - 1
    if min_filter in ("GL_NEAREST_MIPMAP_NEAREST", "GL_LINEAR_MIPMAP_NEAREST"):
        d = int(min(max(np.ceil(lod + 0.5) - 1.0, 0.0), top))
        return one(levels[d], lin)
    if lod >= top:
        return one(levels[top], lin)
    d = int(np.floor(lod))
    (s0, a0), (s1, a1) = one(levels[d], lin), one(levels[d + 1], lin)
    t = F(lod - d)
    return pack(unpack(s0) * (F(1) - t) + unpack(s1) * t), a0 | a1


def quad_pixels(q):
    """Integer pixels (x, row) a screen-aligned quad covers with the
    reference's inclusive edges, and their texture coordinates."""
    X0, R0, X1, R1, su, sv = q[:6]
    s0, t0 = q[6:8] if len(q) > 6 else (0.0, 0.0)
    xs = np.arange(X0, X1 + 1)
    rs = np.arange(R0, R1 + 1)
    X, R = np.meshgrid(xs, rs)
    u = (s0 + su * (X - X0) / (X1 - X0)).astype(F)
    v = (t0 + sv * (R - R0) / (R1 - R0)).astype(F)
    return X, R, u, v


def quad_lod(q, tex_w: int, tex_h: int) -> float:
    X0, R0, X1, R1, su, sv = q[:6]
    return float(np.log2(max(su * tex_w / (X1 - X0), sv * tex_h / (R1 - R0))))


def render(quads, tex: np.ndarray, min_filter: str, mag_filter: str, W: int, H: int, clear=(0, 0, 0, 255)):
    """The expected colour plane (H, W, 4) of lod_quads with GL_REPLACE, a
    (H, W) mask of the pixels whose nearest tap is ambiguous and one of the
    covered pixels."""
    img = np.empty((H, W, 4), np.uint8)
    img[...] = np.array(clear, np.uint8)
    amb = np.zeros((H, W), bool)
    cov = np.zeros((H, W), bool)
    levels = box_chain(tex)
    for q in quads:
        X, R, u, v = quad_pixels(q)
        lod = quad_lod(q, tex.shape[1], tex.shape[0])
        c, a = sample_chain(levels, u, v, lod, min_filter, mag_filter)
        img[R, X] = pack(unpack(c))   # GL_REPLACE: the texel colour, clamped and packed again
        amb[R, X] = a
        cov[R, X] = True
    return img, amb, cov


def compare(got: np.ndarray, want: np.ndarray, amb: np.ndarray, cov: np.ndarray):
    """(largest channel difference outside the ambiguous pixels, fraction of
    the covered pixels that differ there, fraction of the covered pixels that
    are ambiguous)"""
    d = np.abs(got.astype(np.int32) - want.astype(np.int32)).max(-1)
    ok = ~amb
    return int(d[ok].max()), float((d[ok & cov] > 0).mean()), float(amb[cov].mean())
