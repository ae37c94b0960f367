"""An fp64 restatement of the differentiable shading model of
include/mtgl_diff.h for the workload of mytinygl_amd/diff.py, used only to
take finite differences: given the captured per-pixel visibility (G-buffer),
the loss is a function of vertex colours, light colours and texels alone.

The model follows the GL state that diff.draw_view sets and GL's defaults for
the rest: directional lights at eye-space directions light_dir with diffuse
colour light_rgb, ambient (0, 0, 0, 1) and no specular; model ambient
(0.2, 0.2, 0.2, 1); GL_COLOR_MATERIAL in GL_AMBIENT_AND_DIFFUSE mode;
normals through the inverse transpose of the view's modelview.  Its forward
agrees with the oracle's fp32 forward to fp32 rounding (checked in the tests),
which is what makes its finite differences a check of the adjoint.
"""
from __future__ import annotations

import numpy as np

from mytinygl_amd import diff as D

MODEL_AMBIENT = np.array([0.2, 0.2, 0.2])


class Model:
    def __init__(self, params, gbuffers, W: int, H: int):
        self.P = params
        self.W, self.H = W, H
        self.idx = params["idx"].reshape(-1, 3).astype(np.int64)
        self.nvert = params["verts"].shape[0]
        L = params["light_dir"].astype(np.float64)
        self.L = L / np.linalg.norm(L, axis=1, keepdims=True)
        self.views = []
        nrm = params["verts"][:, 3:6].astype(np.float64)
        for v, gb in enumerate(gbuffers):
            mv = D.look_at(params["eyes"][v]).astype(np.float64).reshape(4, 4).T
            nmat = np.linalg.inv(mv[:3, :3]).T
            n = nrm @ nmat.T
            n /= np.linalg.norm(n, axis=1, keepdims=True)
            ndl = np.maximum(n @ self.L.T, 0.0)          # (nvert, lights)
            g = gb.reshape(-1)
            cov = np.nonzero(g["tri"] >= 0)[0]
            self.views.append({
                "ndl": ndl, "cov": cov, "el": self.idx[g["tri"][cov]],
                "b": g["b"][cov].astype(np.float64), "tx": g["texel"][cov].astype(np.int64),
                "fx": g["fx"][cov].astype(np.float64), "fy": g["fy"][cov].astype(np.float64),
                "bg_px": np.nonzero(g["tri"] < 0)[0], "bg": g["texel"][g["tri"] < 0, 0],
            })

    def lit(self, view, vcol, light_rgb):
        """Lit vertex colours (nvert, 4) of one view in fp64."""
        ndl = view["ndl"]
        k = np.clip(vcol.astype(np.float64), 0.0, 1.0)
        s = MODEL_AMBIENT[None, :] + ndl @ light_rgb.astype(np.float64)          # (nvert, 3)
        out = np.empty((self.nvert, 4))
        out[:, :3] = np.clip(k[:, :3] * s, 0.0, 1.0)
        # alpha: the diffuse alpha, plus ambient alpha 1 and diffuse alpha 1 per light
        out[:, 3] = np.clip(k[:, 3] * (1.0 + len(light_rgb) + ndl.sum(1)), 0.0, 1.0)
        return out

    def shade(self, view, vcol, light_rgb, tex, sel=None):
        """Covered pixels of one view (or the subset sel of them) in fp64."""
        lit = self.lit(view, vcol, light_rgb)
        el, b = view["el"], view["b"]
        tx, fx, fy = view["tx"], view["fx"], view["fy"]
        if sel is not None:
            el, b, tx, fx, fy = el[sel], b[sel], tx[sel], fx[sel], fy[sel]
        c = (lit[el] * b[..., None]).sum(1)
        t = tex.reshape(-1, 4).astype(np.float64)
        top = t[tx[:, 0]] * (1 - fx)[:, None] + t[tx[:, 1]] * fx[:, None]
        bot = t[tx[:, 2]] * (1 - fx)[:, None] + t[tx[:, 3]] * fx[:, None]
        tc = top * (1 - fy)[:, None] + bot * fy[:, None]
        return np.clip(c * tc, 0.0, 1.0)

    def images(self, vcol, light_rgb, tex):
        out = np.zeros((len(self.views), self.H * self.W, 4))
        for i, view in enumerate(self.views):
            out[i, view["cov"]] = self.shade(view, vcol, light_rgb, tex)
            bg = view["bg"].astype(np.uint32)
            out[i, view["bg_px"]] = np.stack([(bg >> s) & 0xFF for s in (0, 8, 16, 24)], -1) / 255.0
        return out.reshape(len(self.views), self.H, self.W, 4)

    def loss(self, vcol, light_rgb, tex, target, views=None, sel=None):
        """Sum of squared differences over the covered pixels of the given
        views (all), restricted per view to the pixel subsets sel[i]."""
        tot = 0.0
        for i, view in enumerate(self.views):
            if views is not None and i not in views:
                continue
            s = None if sel is None else sel[i]
            img = self.shade(view, vcol, light_rgb, tex, s)
            cov = view["cov"] if s is None else view["cov"][s]
            tgt = target[i].reshape(-1, 4)[cov].astype(np.float64)
            tot += float(((img - tgt) ** 2).sum())
        return tot

    # pixel subsets a parameter reaches (the rest of the loss does not move)
    def texel_pixels(self, texel):
        return [np.nonzero((v["tx"] == texel).any(1))[0] for v in self.views]

    def vertex_pixels(self, vert):
        return [np.nonzero((v["el"] == vert).any(1))[0] for v in self.views]


def fd_check(model, params, target, g_vcol, g_light, g_tex, rng, n_tex=64, n_vcol=64, n_light=24, h=1e-3,
             h_light=1e-5):
    """Central differences in fp64 on seeded parameter subsets against the
    given gradients; returns the worst relative error and the details.

    The loss is piecewise quadratic in every parameter (the pieces end where a
    lit colour or a pixel reaches the clamp), so a central difference inside
    one piece is the exact derivative.  h = 1e-3 stays inside for texels and
    vertex colours; a light colour moves every vertex it lights at once, and at
    1e-3 a few of the thousands of vertices cross the lighting clamp, so light
    components use h_light."""
    vcol = params["vcol"].astype(np.float64)
    lrgb = params["light_rgb"].astype(np.float64)
    tex = params["tex"].astype(np.float64)
    rows = []
    # texels that some pixel samples, and vertices that some pixel uses
    seen_t = np.unique(np.concatenate([v["tx"].reshape(-1) for v in model.views]))
    seen_v = np.unique(np.concatenate([v["el"].reshape(-1) for v in model.views]))
    for t in rng.choice(seen_t, size=min(n_tex // 4, len(seen_t)), replace=False):
        sel = model.texel_pixels(int(t))
        for ch in range(4):
            tp, tm = tex.copy(), tex.copy()
            tp.reshape(-1, 4)[t, ch] += h
            tm.reshape(-1, 4)[t, ch] -= h
            fd = (model.loss(vcol, lrgb, tp, target, sel=sel) - model.loss(vcol, lrgb, tm, target, sel=sel)) / (2 * h)
            rows.append(("tex", int(t), ch, fd, float(g_tex.reshape(-1, 4)[t, ch])))
    for k in rng.choice(len(seen_v) * 4, size=min(n_vcol, len(seen_v) * 4), replace=False):
        vtx, ch = int(seen_v[k // 4]), int(k % 4)
        sel = model.vertex_pixels(vtx)
        vp, vm = vcol.copy(), vcol.copy()
        vp[vtx, ch] += h
        vm[vtx, ch] -= h
        fd = (model.loss(vp, lrgb, tex, target, sel=sel) - model.loss(vm, lrgb, tex, target, sel=sel)) / (2 * h)
        rows.append(("vcol", vtx, ch, fd, float(g_vcol[vtx, ch])))
    for k in rng.choice(24, size=n_light, replace=False):
        i, ch = int(k // 3), int(k % 3)
        lp, lm = lrgb.copy(), lrgb.copy()
        lp[i, ch] += h_light
        lm[i, ch] -= h_light
        fd = (model.loss(vcol, lp, tex, target) - model.loss(vcol, lm, tex, target)) / (2 * h_light)
        rows.append(("light", i, ch, fd, float(g_light[i, ch])))
    worst = 0.0
    for kind, a, ch, fd, g in rows:
        err = abs(fd - g) / max(1e-3, abs(fd))
        worst = max(worst, err)
    return worst, rows
