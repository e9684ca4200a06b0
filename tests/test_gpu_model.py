"""The model layer of the training step on the GPU.

  * ``activate`` (exp / F.normalize / sigmoid in one HIP launch each way)
    against the torch ops of scene/gaussian_model_cf.py:48-56, values and
    gradients, including zero, sub-eps and huge quaternions and saturated
    scales and opacities;
  * the row learning rate of the fused Adam: one [P,16,3] SH tensor stepped
    with (lead 3, row 48) equals, bit for bit, the same Adam over separate
    f_dc [P,1,3] and f_rest [P,15,3] tensors, and matches torch.optim.Adam
    over those two groups;
  * one training step's gradients through ``model.render`` (fused
    activations, one SH tensor, clamp inside the fused loss) against the
    reference's composition in torch (torch activations, cat of the SH,
    image clamp, then the loss), with the deterministic backward.
"""
import pytest
import torch
import torch.nn.functional as F

from gs_hip import raster, scenes
from gs_hip.loss import fused_photometric_loss
from gs_hip.model import GaussianParams, activate, render
from gs_hip.optim import FusedAdam

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _raw(P, seed):
    g = torch.Generator().manual_seed(seed)
    s = torch.randn((P, 3), generator=g) * 2.0 - 3.0
    r = torch.randn((P, 4), generator=g)
    o = torch.randn((P, 1), generator=g) * 3.0
    # edge rows: zero, sub-eps and huge quaternions; saturating scales and opacities
    r[0] = 0.0
    r[1] = torch.tensor([1e-13, 0.0, -2e-13, 0.0])
    r[2] = torch.tensor([3e18, -1e18, 2e18, 5e17])
    s[3] = torch.tensor([90.0, -110.0, 0.0])
    o[4] = 120.0
    o[5] = -120.0
    return s.to(DEV), r.to(DEV), o.to(DEV)


@pytest.mark.parametrize("P", [1, 257, 10_001])
def test_activate_forward_matches_torch(P):
    s, r, o = (t[:P] for t in _raw(max(P, 6), P))
    sc, ro, op = activate(s, r, o)
    torch.testing.assert_close(sc, torch.exp(s), rtol=2e-7, atol=0)
    torch.testing.assert_close(ro, F.normalize(r), rtol=3e-7, atol=1e-30)
    torch.testing.assert_close(op, torch.sigmoid(o), rtol=2e-7, atol=1e-30)


def test_activate_backward_matches_torch_autograd():
    P = 4099
    s, r, o = _raw(P, 5)
    g = torch.Generator().manual_seed(9)
    gs, gr, go = (torch.randn(t.shape, generator=g).to(DEV) for t in (s, r, o))
    a = [t.clone().requires_grad_(True) for t in (s, r, o)]
    b = [t.clone().requires_grad_(True) for t in (s, r, o)]
    sc, ro, op = activate(*a)
    ((sc * gs).sum() + (ro * gr).sum() + (op * go).sum()).backward()
    tsc, tro, top = torch.exp(b[0]), F.normalize(b[1]), torch.sigmoid(b[2])
    ((tsc * gs).sum() + (tro * gr).sum() + (top * go).sum()).backward()
    torch.testing.assert_close(a[0].grad, b[0].grad, rtol=1e-6, atol=0)
    torch.testing.assert_close(a[2].grad, b[2].grad, rtol=1e-6, atol=1e-30)
    # normalize's backward: per row, relative to the row's gradient scale
    d = (a[1].grad - b[1].grad).abs().amax(dim=1)
    scale = b[1].grad.abs().amax(dim=1).clamp_min(1e-30)
    assert torch.all(d <= 1e-5 * scale), (d / scale).max().item()
    # the zero quaternion: the gradient is g / eps, as torch gives
    torch.testing.assert_close(a[1].grad[0], gr[0] / 1e-12, rtol=1e-6, atol=0)


def test_activate_unused_output_gets_no_gradient():
    s, r, o = (t.clone().requires_grad_(True) for t in _raw(300, 2))
    sc, ro, op = activate(s, r, o)
    (sc.sum() + op.sum()).backward()
    assert r.grad is None
    torch.testing.assert_close(s.grad, torch.exp(s.detach()), rtol=2e-7, atol=0)


@pytest.mark.parametrize("P", [1, 1001, 50_000])
def test_row_adam_equals_two_groups(P):
    g = torch.Generator().manual_seed(P)
    lr_dc, lr_rest = 2.5e-3, 2.5e-3 / 20.0
    init = torch.randn((P, 16, 3), generator=g)
    fused = init.clone().to(DEV).requires_grad_(True)
    dc = init[:, :1].clone().to(DEV).requires_grad_(True)
    rest = init[:, 1:].clone().to(DEV).requires_grad_(True)
    tdc = dc.detach().clone().requires_grad_(True)
    trest = rest.detach().clone().requires_grad_(True)
    o_rows = FusedAdam([{"params": [fused], "lr": lr_rest, "lead_lr": lr_dc, "lead": 3, "row": 48}], eps=1e-15)
    o_two = FusedAdam([{"params": [dc], "lr": lr_dc}, {"params": [rest], "lr": lr_rest}], eps=1e-15)
    o_torch = torch.optim.Adam([{"params": [tdc], "lr": lr_dc}, {"params": [trest], "lr": lr_rest}], lr=0.0,
                               eps=1e-15)
    for _ in range(4):
        gr = torch.randn((P, 16, 3), generator=g).to(DEV)
        fused.grad = gr.clone()
        dc.grad = gr[:, :1].contiguous()
        rest.grad = gr[:, 1:].contiguous()
        tdc.grad = gr[:, :1].contiguous()
        trest.grad = gr[:, 1:].contiguous()
        o_rows.step()
        o_two.step()
        o_torch.step()
    assert torch.equal(fused.detach()[:, :1], dc.detach())
    assert torch.equal(fused.detach()[:, 1:], rest.detach())
    torch.testing.assert_close(fused.detach(), torch.cat([tdc, trest], 1).detach(), rtol=1e-5, atol=1e-7)


def test_row_adam_ragged_rows():
    """row 8 / lead 3 over a length that leaves a partial chunk and a scalar tail."""
    n = 8 * 777
    g = torch.Generator().manual_seed(3)
    x = torch.randn(n, generator=g).to(DEV).requires_grad_(True)
    y = x.detach().clone().requires_grad_(True)
    o1 = FusedAdam([{"params": [x], "lr": 1e-2, "lead_lr": 3e-1, "lead": 3, "row": 8}], eps=1e-15)
    lrs = torch.where(torch.arange(n, device=DEV) % 8 < 3, 3e-1, 1e-2)
    ref = torch.optim.Adam([{"params": [y], "lr": 1.0}], lr=1.0, eps=1e-15)
    for _ in range(3):
        gr = torch.randn(n, generator=g).to(DEV)
        x.grad = gr.clone()
        o1.step()
        # torch Adam with unit lr, scaled per element: p_new = p - lr_i * u
        y0 = y.detach().clone()
        y.grad = gr.clone()
        ref.step()
        with torch.no_grad():
            y.copy_(y0 + (y - y0) * lrs)
    torch.testing.assert_close(x.detach(), y.detach(), rtol=1e-5, atol=1e-6)


def test_training_step_gradients_match_reference_composition():
    sc = scenes.training_scene(seed=4, P=20_000, W=320, H=180)
    cam, bg, target = sc["cam"].to(DEV), sc["bg"].to(DEV), sc["target"].to(DEV)
    raster.set_deterministic(True)
    try:
        gp = GaussianParams(sc["xyz"], sc["features_dc"], sc["features_rest"], sc["scaling"], sc["rotation"],
                            sc["opacity"], sc["sh_degree"], device=DEV)
        out = render(gp, cam, bg, clamp=False)
        loss, _ = fused_photometric_loss(out["image"], target, clamp01=True)
        loss.backward()

        leaves = {k: sc[k].to(DEV).clone().requires_grad_(True)
                  for k in ("xyz", "features_dc", "features_rest", "scaling", "rotation", "opacity")}
        from diff_gaussian_rasterization import GaussianRasterizer
        from gs_hip.camera import raster_settings
        means2D = torch.zeros_like(leaves["xyz"], requires_grad=True) + 0
        means2D.retain_grad()  # scene/gaussian_model_cf.py:793-805
        rast = GaussianRasterizer(raster_settings=raster_settings(cam, bg, sc["sh_degree"]))
        color, _, _ = rast(means3D=leaves["xyz"], means2D=means2D,
                           shs=torch.cat((leaves["features_dc"], leaves["features_rest"]), dim=1),
                           colors_precomp=None, opacities=torch.sigmoid(leaves["opacity"]),
                           scales=torch.exp(leaves["scaling"]), rotations=F.normalize(leaves["rotation"]),
                           cov3D_precomp=None)
        ref_loss, _ = fused_photometric_loss(color.clamp(0, 1), target)
        ref_loss.backward()
    finally:
        raster.set_deterministic(False)
    torch.testing.assert_close(loss, ref_loss, rtol=1e-6, atol=0)

    def close(a, b, name):
        err = (a - b).abs().max().item()
        assert err <= 1e-4 * b.abs().max().item() + 1e-12, (name, err, b.abs().max().item())

    close(gp._xyz.grad, leaves["xyz"].grad, "xyz")
    close(gp._features.grad[:, :1], leaves["features_dc"].grad, "f_dc")
    close(gp._features.grad[:, 1:], leaves["features_rest"].grad, "f_rest")
    close(gp._scaling.grad, leaves["scaling"].grad, "scaling")
    close(gp._rotation.grad, leaves["rotation"].grad, "rotation")
    close(gp._opacity.grad, leaves["opacity"].grad, "opacity")
    # the screen-space leaf carries dL/dmean2D as the reference's retained grad does,
    # and its zero buffer is left untouched for the next step
    close(out["viewspace_points"].grad, means2D.grad, "means2D")
    assert out["viewspace_points"].is_leaf and not torch.any(gp._screen_zeros)
    assert torch.equal(out["visibility_filter"], out["radii"] > 0)
    assert gp.screenspace_leaf() is not out["viewspace_points"]
    assert torch.equal(gp._features_dc, gp._features[:, :1]) and gp._features_rest.shape[1] == 15


def test_adam_in_backward_equals_the_separate_step():
    """render(adam_in_backward=True): the rasteriser's backward applies the
    means' and SH coefficients' Adam step itself (gs_backward_adam) and the
    activations' backward that of the raw scaling, rotation and opacity
    (gs_activate_backward_adam); no parameter gets a gradient.  Over three
    steps of the deterministic backward this equals, bit for bit, backward +
    FusedAdam over every parameter (parameters and both moments)."""
    sc = scenes.training_scene(seed=6, P=30_000, W=320, H=180)
    cam, bg, target = sc["cam"].to(DEV), sc["bg"].to(DEV), sc["target"].to(DEV)
    names = ("_xyz", "_features", "_scaling", "_rotation", "_opacity")
    raster.set_deterministic(True)
    runs = []
    try:
        for fused in (False, True):
            gp = GaussianParams(sc["xyz"], sc["features_dc"], sc["features_rest"], sc["scaling"], sc["rotation"],
                                sc["opacity"], sc["sh_degree"], device=DEV)
            opt = gp.training_setup()
            for _ in range(3):
                out = render(gp, cam, bg, clamp=False, adam_in_backward=fused)
                loss, _ = fused_photometric_loss(out["image"], target, clamp01=True)
                loss.backward()
                if fused:
                    assert all(getattr(gp, n).grad is None for n in names)
                    assert out["viewspace_points"].grad is not None
                opt.step()
                opt.zero_grad(set_to_none=True)
            runs.append((gp, opt))
    finally:
        raster.set_deterministic(False)
    (a, oa), (b, ob) = runs
    for n in names:
        pa, pb = getattr(a, n), getattr(b, n)
        assert torch.equal(pa, pb), n
        sa, sb = oa.state[pa], ob.state[pb]
        assert sa["step"] == sb["step"] == 3, n
        assert torch.equal(sa["exp_avg"], sb["exp_avg"]) and torch.equal(sa["exp_avg_sq"], sb["exp_avg_sq"]), n
    # the parameters did move
    assert not torch.equal(a._xyz, sc["xyz"].to(DEV))


def test_adam_in_backward_refuses_a_pose():
    sc = scenes.training_scene(seed=7, P=2_000, W=64, H=48)
    gp = GaussianParams(sc["xyz"], sc["features_dc"], sc["features_rest"], sc["scaling"], sc["rotation"],
                        sc["opacity"], sc["sh_degree"], device=DEV)
    gp.training_setup()
    pose = torch.tensor([0.0, 0, 0, 0, 0, 0, 1.0], device=DEV)
    with pytest.raises(Exception, match="adam_in_backward"):
        render(gp, sc["cam"].to(DEV), sc["bg"].to(DEV), pose=pose, adam_in_backward=True)
