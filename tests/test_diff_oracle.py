"""The differentiable shading model on the CPU oracle (include/mtgl_diff.h).

- Its forward, given the RGBA8 texture as fp32 (byte / 255), agrees with the
  ordinary render of the same views within one RGBA8 step on every channel
  (the ordinary path re-quantises the bilinear sample to RGBA8), and exactly on
  uncovered pixels.
- The captured G-buffer is well formed.
- An independent fp64 model (diff_model.py) reproduces the oracle's forward to
  fp32 rounding, and its central finite differences reproduce the oracle's
  adjoint (the CPU reference gradient) for texels, vertex colours and all 24
  light components.
"""
import numpy as np
import pytest

from diff_model import Model, fd_check
from mytinygl_amd import diff as D

W = H = 48


@pytest.fixture(scope="module")
def small(oracle):
    P = D.diff_params(grid=16, tex_size=32, views=4)
    _, tq = D.quantise(P["tex"])
    P = dict(P, tex=tq)
    V = D.DiffViews(oracle, P, W, H, keep_frames=True)
    yield P, V
    V.close()


def test_forward_matches_ordinary_render(small):
    P, V = small
    img = V.forward_np(P["vcol"], P["light_rgb"], P["tex"])
    for v in range(V.nviews):
        fr = V.frames[v].view(np.uint8).reshape(H, W, 4).astype(np.int32)
        mine = (np.clip(img[v], 0, 1).astype(np.float32) * np.float32(255.0)).astype(np.int32)
        d = np.abs(fr - mine)
        gb = V.gbuffer(v)
        assert d.max() <= 1, v
        assert not d[gb["tri"] < 0].any(), v
        assert (gb["tri"] >= 0).sum() > 100


def test_gbuffer_well_formed(small):
    P, V = small
    ntri = len(P["idx"]) // 3
    for v in range(V.nviews):
        gb = V.gbuffer(v).reshape(-1)
        cov = gb[gb["tri"] >= 0]
        assert (gb["tri"] >= -1).all() and (cov["tri"] < ntri).all()
        assert np.abs(cov["b"].sum(-1) - 1.0).max() < 1e-5
        assert (cov["b"] >= -1e-6).all()
        assert (cov["texel"] < 32 * 32).all()
        assert ((cov["fx"] >= 0) & (cov["fx"] < 1) & (cov["fy"] >= 0) & (cov["fy"] < 1)).all()


def test_fp64_model_and_finite_differences(small):
    P, V = small
    T = D.perturbed(P, seed=11)
    target = V.forward_np(T["vcol"], T["light_rgb"], T["tex"])
    model = Model(P, [V.gbuffer(v) for v in range(V.nviews)], W, H)
    img = V.forward_np(P["vcol"], P["light_rgb"], P["tex"])
    m_img = model.images(P["vcol"], P["light_rgb"], P["tex"])
    assert np.abs(m_img - img).max() < 2e-6
    grad = 2.0 * (img - target)
    gv, gl, gt = V.backward_np(P["vcol"], P["light_rgb"], P["tex"], grad)
    worst, rows = fd_check(model, P, target, gv, gl, gt, np.random.default_rng(5), n_tex=64, n_vcol=64, n_light=24)
    assert len(rows) == 64 + 64 + 24
    assert worst < 2e-3, [r for r in rows if abs(r[3] - r[4]) > 2e-3 * max(1e-3, abs(r[3]))][:5]
    assert np.abs(gl).sum() > 0 and np.abs(gt).sum() > 0 and np.abs(gv).sum() > 0


def test_wrong_shapes_raise_before_any_call(small):
    """Arguments whose shapes differ from the captured views' are refused
    with ValueError (the kernels would read or write past them)."""
    P, V = small
    img = V.forward_np(P["vcol"], P["light_rgb"], P["tex"])
    bad = {
        "vcol": P["vcol"][:-1],
        "light_rgb": P["light_rgb"][:4],
        "tex": P["tex"][:, :16],
    }
    for name, arr in bad.items():
        args = dict(vcol=P["vcol"], light_rgb=P["light_rgb"], tex=P["tex"])
        args[name] = arr
        with pytest.raises(ValueError, match=name):
            V.forward_np(**args)
        with pytest.raises(ValueError, match=name):
            V.backward_np(grad=img, **args)
    with pytest.raises(ValueError, match="grad"):
        V.backward_np(P["vcol"], P["light_rgb"], P["tex"], img[:1])
    # float64 inputs of the right shapes are converted, not misread
    a = V.forward_np(P["vcol"].astype(np.float64), P["light_rgb"], P["tex"])
    assert np.array_equal(a, img)


def check_unarmed_draw_is_refused(g):
    """A draw made before mtgl_diff_arm left no last-writer plane: capturing it
    is refused with MTGL_ERR_UNSUPPORTED (4); after arming it is taken."""
    import ctypes as C
    P = D.diff_params(grid=8, tex_size=16, views=1)
    tex8, _ = D.quantise(P["tex"])
    D._bind(g.lib)
    h = C.c_void_p()
    assert g.lib.mtgl_diff_create(P["verts"].shape[0], 16, 16, C.byref(h)) == 0
    try:
        ctx = g.create_context(W, H)
        try:
            D.draw_view(g, P, 0, W, H, tex8)
            assert g.lib.mtgl_diff_add_view(h, g.device(ctx)) == 4
            assert g.lib.mtgl_diff_arm(g.device(ctx)) == 0
            D.draw_view(g, P, 0, W, H, tex8)
            assert g.lib.mtgl_diff_add_view(h, g.device(ctx)) == 0
        finally:
            g.destroy_context(ctx)
            g.release()
    finally:
        g.lib.mtgl_diff_destroy(h)


def test_unarmed_draw_is_refused(oracle):
    check_unarmed_draw_is_refused(oracle)
