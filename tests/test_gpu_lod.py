"""The full-mip-chain extension (MTGL_LOD_FULL_CHAIN) on the device against
the fp32/fp64 model of tests/lod_model.py (pinned against the oracle in
tests/test_lod_model.py): screen-aligned quads whose LODs run from
magnification to beyond the last level, under every mipmapped minification
filter.  Error budget, stated here because the extension has no reference
to be bit-exact to: outside the pixels whose nearest tap lies on a texel edge,
at most one RGBA8 step on at most 5% of the covered pixels and none on the
rest.  In the default mode the same scene renders the reference's sampling
(the oracle's image)."""
import numpy as np
import pytest

import lod_model as M
from conftest import plane_diffs
from mytinygl_amd import scenes as S

pytestmark = pytest.mark.gpu

W = H = 256
FILTERS = ["GL_NEAREST_MIPMAP_NEAREST", "GL_LINEAR_MIPMAP_NEAREST", "GL_NEAREST_MIPMAP_LINEAR",
           "GL_LINEAR_MIPMAP_LINEAR"]


def _render(g, tex, minf, mode):
    def fn(gg, W_, H_, seed):
        gg.texture_lod_mode(gg.gl_get_current_context(), mode)
        S.lod_quads(gg, W_, H_, tex, minf, "GL_LINEAR")
    planes, err = S.render(g, fn, W, H, 0)
    assert err == 0
    c = planes[0]
    return planes, np.stack([(c >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.uint8)


@pytest.mark.parametrize("minf", FILTERS)
@pytest.mark.parametrize("seed", [1, 2])
def test_full_chain_vs_model(mtgl, minf, seed):
    tex = S.random_rgba(64, 64, seed)
    _, got = _render(mtgl, tex, minf, 1)
    want, amb, cov = M.render(S.LOD_QUADS, tex, minf, "GL_LINEAR", W, H)
    worst, frac, amb_frac = M.compare(got, want, amb, cov)
    print(f"{minf}/{seed}: worst {worst}, differing {frac:.4f}, ambiguous {amb_frac:.4f}", flush=True)
    assert worst <= 1 and frac < 0.05 and amb_frac < 0.1


def test_reference_mode_is_the_oracle(mtgl, oracle):
    tex = S.random_rgba(64, 64, 4)
    a, _ = _render(oracle, tex, "GL_LINEAR_MIPMAP_LINEAR", 0)
    b, _ = _render(mtgl, tex, "GL_LINEAR_MIPMAP_LINEAR", 0)
    assert plane_diffs(a, b) == {"color": 0, "depth": 0, "stencil": 0}


def test_full_chain_differs_from_reference(mtgl):
    """Past level 1 the chain samples other texels than the reference's two
    levels."""
    tex = S.random_rgba(64, 64, 5)
    _, a = _render(mtgl, tex, "GL_LINEAR_MIPMAP_LINEAR", 0)
    _, b = _render(mtgl, tex, "GL_LINEAR_MIPMAP_LINEAR", 1)
    assert (a != b).any(-1).sum() > 1000
