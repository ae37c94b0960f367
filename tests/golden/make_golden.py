#!/usr/bin/env python3
"""Writes tests/golden/golden.json from the reference renderer.

Run where the reference sources are present and oracle/_ref/libmtgl_ref.so
has been built (make -C oracle ref): the reference compiled strict from its
own sources (no FMA contraction, no fast-math).  The fixture holds data only:

  api      observable results of the GL call scripts in tests/api_scripts.py
  scenes   SHA-256 of the colour, depth and stencil planes and the final
           glGetError of every feature scene of mytinygl_amd/scenes.py at
           the parity-test size, seeds 1-3 (scenes whose reference behaviour
           is undefined are left out)
  configs  the same for the icosphere 640x480, textured sphere 1920x1080 and
           overdraw 3840x2160 workloads at full size, plus their fragment
           counts: the number of (triangle, pixel) pairs that pass the
           reference's inside test, measured by redrawing the workload with
           the depth test off and stencil GL_ALWAYS / GL_INCR on every
           fragment, then summing the stencil plane (exact while no pixel
           reaches 255, which the script checks)

The fixture pins the CPU oracle (tests/test_oracle_golden.py) on machines
where the reference itself is absent.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from mytinygl_amd import scenes as S  # noqa: E402
from mytinygl_amd.gl import ENUMS as E, GLLibrary  # noqa: E402
import api_scripts  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
OUT = os.path.join(HERE, "golden.json")
SEEDS = (1, 2, 3)

CONFIGS = {
    "icosphere_640x480": (S.setup_icosphere, S.draw_icosphere, 640, 480),
    "textured_sphere_1920x1080": (S.setup_textured_sphere, S.draw_textured_sphere, 1920, 1080),
    "overdraw_3840x2160": (S.setup_overdraw, S.draw_overdraw, 3840, 2160),
    "libm_spheres_1920x1080": (S.setup_libm_spheres, S.draw_libm_spheres, 1920, 1080),
}


def scene_size(name):
    return (160, 120) if "small" in name else (96, 72)


def digest(planes, err):
    c, d, s = planes
    return {"color": hashlib.sha256(c.tobytes()).hexdigest(), "depth": hashlib.sha256(d.tobytes()).hexdigest(),
            "stencil": hashlib.sha256(s.tobytes()).hexdigest(), "error": int(err)}


def render_config(g, setup, draw, W, H):
    ctx = g.create_context(W, H)
    try:
        g.glClearColor(0.0, 0.0, 0.0, 1.0)
        g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
        n = setup(g, W, H)
        draw(g, n)
        planes = g.read_framebuffer(ctx, W, H)
        err = g.glGetError()
    finally:
        g.destroy_context(ctx)
        g.release()
    return planes, err


def count_fragments(g, setup, draw, W, H):
    ctx = g.create_context(W, H)
    try:
        g.glClearStencil(0)
        g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"])
        n = setup(g, W, H)
        g.glDisable(E["GL_DEPTH_TEST"])
        g.glEnable(E["GL_STENCIL_TEST"])
        g.glStencilFunc(E["GL_ALWAYS"], 0, 0xFF)
        g.glStencilOp(E["GL_INCR"], E["GL_INCR"], E["GL_INCR"])
        g.glStencilMask(0xFF)
        draw(g, n)
        _, _, st = g.read_framebuffer(ctx, W, H)
    finally:
        g.destroy_context(ctx)
        g.release()
    if int(st.max()) >= 255:
        return None
    return int(st.astype(np.int64).sum())


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is not built: make -C oracle ref (needs the reference sources)")
    g = GLLibrary(REF, "reference")
    out = {"generator": "tests/golden/make_golden.py", "library": "oracle/_ref/libmtgl_ref.so (strict build)",
           "api": {}, "scenes": {}, "configs": {}}
    for name in api_scripts.SCRIPTS:
        out["api"][name] = json.loads(json.dumps(api_scripts.run_script(g, name)))
    for name, fn in sorted(S.FEATURE_SCENES.items()):
        if name in S.UNDEFINED_IN_REFERENCE:
            continue
        W, H = scene_size(name)
        for seed in SEEDS:
            planes, err = S.render(g, fn, W, H, seed)
            out["scenes"][f"{name}/{seed}"] = dict(size=[W, H], **digest(planes, err))
        print("scene", name, flush=True)
    for name, (setup, draw, W, H) in CONFIGS.items():
        planes, err = render_config(g, setup, draw, W, H)
        rec = dict(size=[W, H], **digest(planes, err))
        rec["fragments"] = count_fragments(g, setup, draw, W, H)
        out["configs"][name] = rec
        print("config", name, rec["fragments"], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
