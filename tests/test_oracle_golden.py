"""The CPU oracle against fixtures recorded from the reference renderer
(tests/golden/golden.json, written by tests/golden/make_golden.py from the
strict build of the reference's own sources).  Bit-exact planes, identical
glGetError results and GL state query results, and the fragment count of
each full-size workload."""
import hashlib
import json
import os

import pytest

from conftest import ROOT
from mytinygl_amd import scenes as S
import api_scripts

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
CONFIGS = {
    "icosphere_640x480": (S.setup_icosphere, S.draw_icosphere),
    "textured_sphere_1920x1080": (S.setup_textured_sphere, S.draw_textured_sphere),
    "overdraw_3840x2160": (S.setup_overdraw, S.draw_overdraw),
    "libm_spheres_1920x1080": (S.setup_libm_spheres, S.draw_libm_spheres),
}


def _digest(planes):
    return {k: hashlib.sha256(p.tobytes()).hexdigest() for k, p in zip(("color", "depth", "stencil"), planes)}


@pytest.mark.parametrize("name", sorted(GOLDEN["api"]))
def test_api_script(oracle, name):
    got = json.loads(json.dumps(api_scripts.run_script(oracle, name)))
    assert got == GOLDEN["api"][name]


@pytest.mark.parametrize("key", sorted(GOLDEN["scenes"]))
def test_scene(oracle, key):
    name, seed = key.split("/")
    rec = GOLDEN["scenes"][key]
    W, H = rec["size"]
    planes, err = S.render(oracle, S.FEATURE_SCENES[name], W, H, int(seed))
    assert err == rec["error"]
    got = _digest(planes)
    for plane in ("depth", "stencil", "color"):
        assert got[plane] == rec[plane], f"{key}: {plane} plane differs from the reference"


@pytest.mark.parametrize("cfg", sorted(GOLDEN["configs"]))
def test_config_full_size(oracle, cfg):
    rec = GOLDEN["configs"][cfg]
    W, H = rec["size"]
    setup, draw = CONFIGS[cfg]
    ctx = oracle.create_context(W, H)
    try:
        oracle.glClearColor(0.0, 0.0, 0.0, 1.0)
        oracle.glClear(S.E["GL_COLOR_BUFFER_BIT"] | S.E["GL_DEPTH_BUFFER_BIT"] | S.E["GL_STENCIL_BUFFER_BIT"])
        n = setup(oracle, W, H)
        oracle.stats_reset(ctx)
        draw(oracle, n)
        planes = oracle.read_framebuffer(ctx, W, H)
        st = oracle.stats(ctx)
        err = oracle.glGetError()
    finally:
        oracle.destroy_context(ctx)
        oracle.release()
    assert err == rec["error"]
    got = _digest(planes)
    for plane in ("depth", "stencil", "color"):
        assert got[plane] == rec[plane], f"{cfg}: {plane} plane differs from the reference"
    if rec["fragments"] is not None:
        assert st.fragments == rec["fragments"]
