"""Pins the fp64/fp32 model of the full-mip-chain extension (tests/lod_model.py)
against the CPU oracle where the reference's own sampling coincides with it:
magnified quads (level 0 with the magnification filter in both), and quads at
exactly 2x minification under GL_NEAREST_MIPMAP_NEAREST and
GL_LINEAR_MIPMAP_NEAREST, where the reference's per-triangle LOD is exactly 1
(level 1) and the chain's nearest level is ceil(1.5) - 1 = 1.  The model's
texture coordinates are exact where the oracle's come from fp32 weights, so
a bilinear sample may truncate to a neighbouring byte: the bar is at most one
RGBA8 step on at most 5% of the covered pixels, none on the rest, outside
the pixels whose nearest tap lies on a texel edge.  The box-filter
chain itself is checked against a direct mean of the 2^k x 2^k block for a
texture whose blocks are constant."""
import numpy as np
import pytest

import lod_model as M
from conftest import plane_diffs
from mytinygl_amd import scenes as S

W = H = 256


def _tex(seed, size=64):
    return S.random_rgba(size, size, seed)


def _oracle_quads(g, tex, minf, magf, quads):
    def fn(gg, W_, H_, seed):
        S.lod_quads(gg, W_, H_, tex, minf, magf, quads)
    (planes, err) = S.render(g, fn, W, H, 0)
    assert err == 0
    c = planes[0]
    return np.stack([(c >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.uint8)


MAGNIFIED = [(4, 4, 67, 67, 0.5, 0.5), (72, 4, 199, 51, 1.3, 0.4), (4, 72, 100, 250, 0.2, 0.9)]


@pytest.mark.parametrize("magf", ["GL_NEAREST", "GL_LINEAR"])
@pytest.mark.parametrize("seed", [1, 2])
def test_model_magnified_equals_oracle(oracle, magf, seed):
    tex = _tex(seed)
    got = _oracle_quads(oracle, tex, "GL_LINEAR_MIPMAP_LINEAR", magf, MAGNIFIED)
    want, amb, cov = M.render(MAGNIFIED, tex, "GL_LINEAR_MIPMAP_LINEAR", magf, W, H)
    worst, frac, amb_frac = M.compare(got, want, amb, cov)
    assert worst <= 1 and frac < 0.05 and amb_frac < 0.1


@pytest.mark.parametrize("minf", ["GL_NEAREST_MIPMAP_NEAREST", "GL_LINEAR_MIPMAP_NEAREST"])
def test_model_level1_equals_oracle(oracle, minf):
    tex = _tex(3)
    # 2 texels per pixel: the reference's per-triangle LOD is 0.5 log2(4) = 1
    # (the second quad: 0.5 log2(2.02) >= 0.5), the chain's 1 (1.01); taps
    # start a quarter of a level-1 texel in, away from texel edges
    o = 0.25 / 32
    quads = [(4, 4, 67, 67, 2.0 * 63 / 64, 2.0 * 63 / 64, o, o),
             (108, 4, 171, 131, 127 / 64, 2.0 * 127 / 128, o, o)]
    assert M.quad_lod(quads[0], 64, 64) == 1.0 and 1.0 < M.quad_lod(quads[1], 64, 64) < 1.5
    got = _oracle_quads(oracle, tex, minf, "GL_LINEAR", quads)
    want, amb, cov = M.render(quads, tex, minf, "GL_LINEAR", W, H)
    worst, frac, amb_frac = M.compare(got, want, amb, cov)
    assert worst <= 1 and frac < 0.05 and amb_frac < 0.1


def test_chain_is_block_mean():
    """A texture constant on 8x8 blocks: level k (k <= 3) equals the block
    values exactly, and the chain stops at a side of 1."""
    rng = np.random.default_rng(5)
    blocks = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    tex = np.repeat(np.repeat(blocks, 8, 0), 8, 1)
    lv = M.box_chain(tex)
    assert [l.shape[0] for l in lv] == [64, 32, 16, 8, 4, 2, 1]
    for k in range(4):
        f = 8 >> k
        assert np.array_equal(lv[k][::f, ::f], blocks) if f >= 1 else True
    assert np.array_equal(lv[3], blocks)
