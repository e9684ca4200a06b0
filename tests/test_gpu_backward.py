"""Backward parity of the HIP rasteriser against the fp32 CPU oracle's analytic
gradients (which are themselves pinned by fp64 central differences in
tests/test_oracle_gradients.py).

Tolerance: |gpu - oracle| <= 2e-4 * max|oracle| + 2e-3 * |oracle| per entry.
Both sides recompute the same alphas bit-exactly; the difference is only the
summation order of the per-pixel terms (wave reductions + float atomics vs
per-thread buffers).
"""
import numpy as np
import pytest
import torch

from gpu_helpers import DEV, assert_close_grad, np_, run_gpu, run_oracle
from gs_hip import raster, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cotangents(H, W, seed):
    g = torch.Generator().manual_seed(seed)
    dc = (torch.rand((3, H, W), generator=g) - 0.5) / (H * W)
    dd = (torch.rand((1, H, W), generator=g) - 0.5) / (H * W)
    da = (torch.rand((1, H, W), generator=g) - 0.5) / (H * W)
    return dc, dd, da


def _check(sc, deg, seed=0, deterministic=None, **kw):
    g = run_gpu(sc, deg, **kw)
    scene, cam, o = run_oracle(sc, deg, **{k: v for k, v in kw.items() if k != "binning"})
    H, W = cam["H"], cam["W"]
    dc, dd, da = _cotangents(H, W, seed)
    og = O.backward(scene, cam, o, dc.numpy(), dd.numpy(), da.numpy())
    means, shs, col, op, sc_, ro, cv, po = g["inputs"]
    gg = raster.backward(g["rs"], g["st"], means, shs, col, op, sc_, ro, cv, po, dc.to(DEV), dd.to(DEV),
                         da.to(DEV), deterministic=deterministic)
    torch.cuda.synchronize()
    vis = o["radii"] > 0
    assert_close_grad("means2D", np_(gg["means2D"]), og["means2D"])
    assert_close_grad("opacities", np_(gg["opacities"]), og["opacities"])
    assert_close_grad("colors", np_(gg["colors"]), og["colors"])
    assert_close_grad("means3D", np_(gg["means3D"]), og["means3D"])
    assert_close_grad("cov3D", np_(gg["cov3D"])[vis], og["cov3D"][vis])
    if gg["shs"] is not None:
        assert_close_grad("shs", np_(gg["shs"]), og["shs"])
    if gg["scales"] is not None:
        assert_close_grad("scales", np_(gg["scales"]), og["scales"])
        assert_close_grad("rotations", np_(gg["rotations"]), og["rotations"])
    if gg["pose"] is not None:
        assert_close_grad("pose", np_(gg["pose"]), og["pose7"], rtol=5e-3, atol_frac=1e-3)
    if deterministic:
        # fixed-order reductions: a second backward of the same forward is bit-identical
        again = raster.backward(g["rs"], g["st"], means, shs, col, op, sc_, ro, cv, po, dc.to(DEV), dd.to(DEV),
                                da.to(DEV), deterministic=True)
        for k, v in gg.items():
            if v is not None:
                assert torch.equal(v, again[k]), k
    return gg, og


@pytest.mark.parametrize("binning", [0, 1])
def test_deterministic_backward(binning):
    """Fixed-order mode: oracle parity at the same tolerances, bitwise repeatable."""
    sc = scenes.plain_scene(seed=40 + binning, P=6000)
    pose = scenes.random_pose(9, 0.05, 0.05)
    _check(sc, 3, deterministic=True, pose=pose, binning=binning)


def test_plain_scene_backward():
    sc = scenes.plain_scene(seed=0)
    _check(sc, 3)


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_backward_sh_degrees(deg):
    sc = scenes.plain_scene(seed=20 + deg, P=3000)
    _check(sc, deg, seed=deg)


def test_backward_colors_and_cov3D_precomp():
    sc = scenes.plain_scene(seed=30, P=3000)
    colors = torch.rand((3000, 3), generator=torch.Generator().manual_seed(1))
    _check(sc, 3, colors=colors)
    cov = torch.zeros((3000, 6))
    s2 = sc["scales"] ** 2
    cov[:, 0], cov[:, 3], cov[:, 5] = s2[:, 0], s2[:, 1], s2[:, 2]
    _check(sc, 3, cov3D=cov)


def test_backward_with_pose():
    sc = scenes.plain_scene(seed=31, P=4000)
    pose = scenes.random_pose(5, 0.08, 0.08)
    gg, og = _check(sc, 3, pose=pose)
    assert np.abs(og["pose7"]).max() > 0


def test_backward_white_background_ragged():
    sc = scenes.plain_scene(seed=32, P=4000, W=200, H=120)
    sc["bg"] = torch.tensor([1.0, 0.5, 0.25])
    _check(sc, 1)


def test_autograd_function_end_to_end():
    """The autograd wrapper routes every gradient to the right input, means2D
    receives the screen-space gradient and radii are non-differentiable."""
    from diff_gaussian_rasterization import GaussianRasterizer
    from gs_hip import camera as gcam
    sc = scenes.plain_scene(seed=33, P=2000)
    cam = sc["cam"].to(DEV)
    rs = gcam.raster_settings(cam, sc["bg"].to(DEV), 3)
    means = sc["means"].to(DEV).requires_grad_(True)
    means2D = torch.zeros_like(means, requires_grad=True)
    shs = sc["shs"].to(DEV).requires_grad_(True)
    op = sc["opacities"].to(DEV).requires_grad_(True)
    scl = sc["scales"].to(DEV).requires_grad_(True)
    rot = sc["rotations"].to(DEV).requires_grad_(True)
    color, radii, depth = GaussianRasterizer(rs)(means3D=means, means2D=means2D, shs=shs, colors_precomp=None,
                                                 opacities=op, scales=scl, rotations=rot, cov3D_precomp=None)
    assert radii.dtype == torch.int32 and not radii.requires_grad
    tgt = sc["target"].to(DEV)
    loss = (color.clamp(0, 1) - tgt).abs().mean() + 0.1 * depth.mean()
    loss.backward()
    for t in (means, means2D, shs, op, scl, rot):
        assert t.grad is not None and torch.isfinite(t.grad).all()
    assert float(means2D.grad[:, :2].abs().sum()) > 0
    assert float(means2D.grad[:, 2].abs().sum()) == 0
