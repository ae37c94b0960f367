"""GPU parity of the differentiable path (include/mtgl_diff.h) behind the
torch.autograd.Function of mytinygl_amd/diff.py.

- G-buffer and forward images bit-identical to the CPU oracle's;
- adjoint (fp32 atomics) against the oracle's CPU reference gradient
  (accumulated in double) and against fp64 central differences of an
  independent model (diff_model.py) on seeded subsets: 64 texel components,
  64 vertex-colour components and all 24 light components;
- the full workload (64 views of 256x256, 100,352 triangles, 1024x1024 fp32
  texture, 8 lights): forward within one RGBA8 step of the ordinary renders,
  and finite differences on a subset.
"""
import numpy as np
import pytest

from diff_model import Model, fd_check
from mytinygl_amd import diff as D

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MW = MH = 96


@pytest.fixture(scope="module")
def med(mtgl, oracle):
    P = D.diff_params(grid=48, tex_size=64, views=8)
    _, tq = D.quantise(P["tex"])
    P = dict(P, tex=tq)
    Vd = D.DiffViews(mtgl, P, MW, MH)
    Vo = D.DiffViews(oracle, P, MW, MH)
    yield P, Vd, Vo
    Vd.close()
    Vo.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


def test_gbuffer_matches_oracle(med):
    P, Vd, Vo = med
    for v in range(Vd.nviews):
        a, b = Vd.gbuffer(v), Vo.gbuffer(v)
        assert a.tobytes() == b.tobytes(), f"view {v}: {(a != b).sum()} pixels differ"


def test_forward_matches_oracle(med):
    P, Vd, Vo = med
    ref = Vo.forward_np(P["vcol"], P["light_rgb"], P["tex"])
    img = D.render(_dev(P["vcol"]), _dev(P["light_rgb"]), _dev(P["tex"]), Vd).cpu().numpy()
    print(f"forward: max |device - oracle| = {np.abs(img - ref).max()}")
    assert np.array_equal(img, ref)


def _grads(P, Vd, target):
    vc, lr, tx = (_dev(P[k]).requires_grad_() for k in ("vcol", "light_rgb", "tex"))
    img = D.render(vc, lr, tx, Vd)
    loss = ((img - _dev(target)) ** 2).sum()
    loss.backward()
    return img.detach().cpu().numpy(), vc.grad.cpu().numpy(), lr.grad.cpu().numpy(), tx.grad.cpu().numpy()


def test_backward_matches_reference_and_fd64(med):
    P, Vd, Vo = med
    T = D.perturbed(P, seed=12)
    target = Vo.forward_np(T["vcol"], T["light_rgb"], T["tex"])
    img, gv, gl, gt = _grads(P, Vd, target)
    rv, rl, rt = Vo.backward_np(P["vcol"], P["light_rgb"], P["tex"], 2.0 * (img - target))
    for name, a, b in (("vcol", gv, rv), ("light", gl, rl), ("tex", gt, rt)):
        err = np.abs(a - b).max() / max(1e-12, np.abs(b).max())
        print(f"{name}: max |device - reference| / max |reference| = {err:.2e}")
        assert err < 1e-5, name
    model = Model(P, [Vd.gbuffer(v) for v in range(Vd.nviews)], MW, MH)
    worst, rows = fd_check(model, P, target, gv, gl, gt, np.random.default_rng(7))
    print(f"fd64: {len(rows)} components, worst relative error {worst:.2e}")
    assert len(rows) == 64 + 64 + 24
    assert worst < 2e-3


def test_full_workload(mtgl):
    P = D.diff_params()
    _, tq = D.quantise(P["tex"])
    Vd = D.DiffViews(mtgl, P, 256, 256, keep_frames=True)
    try:
        assert Vd.nviews == 64 and Vd.nvert == 225 * 225 and len(P["idx"]) == 3 * 100352
        img = D.render(_dev(P["vcol"]), _dev(P["light_rgb"]), _dev(tq), Vd).cpu().numpy()
        worst = 0
        for v in range(Vd.nviews):
            fr = Vd.frames[v].view(np.uint8).reshape(256, 256, 4).astype(np.int32)
            mine = (np.clip(img[v], 0, 1).astype(np.float32) * np.float32(255.0)).astype(np.int32)
            worst = max(worst, int(np.abs(fr - mine).max()))
        print(f"full: max RGBA8 difference to the ordinary renders {worst}")
        assert worst <= 1
        T = D.perturbed(P, seed=13)
        target = D.render(_dev(T["vcol"]), _dev(T["light_rgb"]), _dev(T["tex"]), Vd).cpu().numpy()
        img, gv, gl, gt = _grads(P, Vd, target)
        assert np.isfinite(gv).all() and np.isfinite(gl).all() and np.isfinite(gt).all()
        model = Model(P, [Vd.gbuffer(v) for v in range(Vd.nviews)], 256, 256)
        worst, rows = fd_check(model, P, target, gv, gl, gt, np.random.default_rng(8), n_tex=32, n_vcol=16, n_light=6)
        print(f"full fd64: {len(rows)} components, worst relative error {worst:.2e}")
        assert worst < 2e-3
    finally:
        Vd.close()


def test_unarmed_draw_is_refused_on_device(mtgl):
    from test_diff_oracle import check_unarmed_draw_is_refused
    check_unarmed_draw_is_refused(mtgl)
