"""Parity at the workloads' own sizes.

* Training step, 500,000 Gaussians, 960x540, SH degree 3: the render backward
  against the fp32 oracle's gradients for every Gaussian, in both binning
  schedules and with the float-atomic and the fixed-order reductions.  At this
  size the sorted list has about 5.7M entries, tiles run several 256-entry
  batches, and the chunked backward resumes from the forward's checkpoints in
  longest-first order -- none of which the small scenes reach.  The cotangent
  is the training step's: a colour gradient only, no depth or alpha term.
* The same step with the Adam update inside the backward equals the separate
  ``backward(); step()`` bit for bit (fixed-order reductions), and within
  the gradient tolerance with float atomics.

The oracle's integers (radii, list length) must match exactly; gradients use
the tolerance of tests/gpu_helpers.assert_close_grad.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from gpu_helpers import DEV, assert_close_grad, np_
from gs_hip import _lib, raster, scenes
from gs_hip import camera as gcam
from gs_hip.loss import fused_photometric_loss
from gs_hip.model import GaussianParams, render
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P_C2, W_C2, H_C2 = 500_000, 960, 540


@pytest.fixture(scope="module")
def c2():
    """The training scene's activated tensors, the oracle forward and its
    backward under a seeded colour cotangent (and under a depth one too)."""
    print("C2 oracle forward + backward", flush=True)
    sc = scenes.training_scene(seed=0, P=P_C2, W=W_C2, H=H_C2)
    shs = torch.cat([sc["features_dc"], sc["features_rest"]], 1).contiguous()
    t = dict(means=sc["xyz"], shs=shs, opacities=torch.sigmoid(sc["opacity"]),
             scales=torch.exp(sc["scaling"]), rotations=F.normalize(sc["rotation"]))
    cam = gcam.oracle_camera(sc["cam"], sc["bg"], 3)
    scene = {k: v.numpy() for k, v in t.items()}
    o = O.forward(scene, cam)
    g = torch.Generator().manual_seed(11)
    dc = (torch.rand((3, H_C2, W_C2), generator=g) - 0.5) / (H_C2 * W_C2)
    dd = (torch.rand((1, H_C2, W_C2), generator=g) - 0.5) / (H_C2 * W_C2)
    og = O.backward(scene, cam, o, dc.numpy())
    ogd = O.backward(scene, cam, o, dc.numpy(), dd.numpy())
    return dict(sc=sc, t=t, o=o, og=og, ogd=ogd, dc=dc, dd=dd)


def _gpu(c2, binning, deterministic, with_depth=False):
    sc, t = c2["sc"], c2["t"]
    rs = gcam.raster_settings(sc["cam"].to(DEV), sc["bg"].to(DEV), 3)
    args = [t["means"].to(DEV), t["shs"].to(DEV), None, t["opacities"].to(DEV), t["scales"].to(DEV),
            t["rotations"].to(DEV), None]
    color, radii, depth, alpha, st = raster.forward(rs, *args, binning=binning)
    dd = c2["dd"].to(DEV) if with_depth else None
    gg = raster.backward(rs, st, *args, None, c2["dc"].to(DEV), dd, None, deterministic=deterministic)
    torch.cuda.synchronize()
    return color, radii, st, gg


def _compare(gg, og, vis):
    assert_close_grad("means2D", np_(gg["means2D"]), og["means2D"])
    assert_close_grad("opacities", np_(gg["opacities"]), og["opacities"])
    assert_close_grad("colors", np_(gg["colors"]), og["colors"])
    assert_close_grad("cov3D", np_(gg["cov3D"])[vis], og["cov3D"][vis])
    assert_close_grad("means3D", np_(gg["means3D"]), og["means3D"])
    assert_close_grad("shs", np_(gg["shs"]), og["shs"])
    assert_close_grad("scales", np_(gg["scales"]), og["scales"])
    assert_close_grad("rotations", np_(gg["rotations"]), og["rotations"])


@pytest.mark.parametrize("binning", [_lib.BINNING_DEPTH_FIRST, _lib.BINNING_SORT64])
@pytest.mark.parametrize("deterministic", [False, True])
def test_training_step_backward_matches_oracle(c2, binning, deterministic):
    print("C2 backward: binning", binning, "deterministic", deterministic, flush=True)
    o, og = c2["o"], c2["og"]
    color, radii, st, gg = _gpu(c2, binning, deterministic)
    assert st.num_rendered == o["num_rendered"] and st.num_rendered > 5_000_000
    assert np.array_equal(np_(radii), o["radii"])
    assert float(np.abs(np_(color) - o["color"]).max()) <= 1e-5
    _compare(gg, og, o["radii"] > 0)
    if deterministic:
        again = _gpu(c2, binning, True)[3]
        for k, v in gg.items():
            if v is not None:
                assert torch.equal(v, again[k]), k


def test_training_step_backward_with_depth_cotangent(c2):
    print("C2 backward with a depth cotangent", flush=True)
    o = c2["o"]
    _, _, _, gg = _gpu(c2, _lib.BINNING_DEPTH_FIRST, False, with_depth=True)
    _compare(gg, c2["ogd"], o["radii"] > 0)


def _train(sc, steps, fused, deterministic):
    raster.set_deterministic(deterministic)
    try:
        gp = GaussianParams(sc["xyz"], sc["features_dc"], sc["features_rest"], sc["scaling"], sc["rotation"],
                            sc["opacity"], sc["sh_degree"], device=DEV)
        opt = gp.training_setup()
        cam, bg, target = sc["cam"].to(DEV), sc["bg"].to(DEV), sc["target"].to(DEV)
        for _ in range(steps):
            out = render(gp, cam, bg, clamp=False, adam_in_backward=fused)
            loss, _ = fused_photometric_loss(out["image"], target, clamp01=True)
            loss.backward()
            opt.step()
            opt.zero_grad(set_to_none=True)
        torch.cuda.synchronize()
    finally:
        raster.set_deterministic(False)
    return gp, opt


NAMES = ("_xyz", "_features", "_scaling", "_rotation", "_opacity")


def test_adam_in_backward_at_full_size_is_bitwise_the_separate_step(c2):
    """Fixed-order reductions: two training steps at 500k / 960x540 with the
    Adam step inside both backward kernels equal backward + FusedAdam bit for
    bit (parameters and both moments)."""
    print("C2 adam-in-backward, deterministic", flush=True)
    sc = c2["sc"]
    a, oa = _train(sc, 2, False, True)
    b, ob = _train(sc, 2, True, True)
    for n in NAMES:
        pa, pb = getattr(a, n), getattr(b, n)
        assert torch.equal(pa, pb), n
        sa, sb = oa.state[pa], ob.state[pb]
        assert sa["step"] == sb["step"] == 2, n
        assert torch.equal(sa["exp_avg"], sb["exp_avg"]), n
        assert torch.equal(sa["exp_avg_sq"], sb["exp_avg_sq"]), n
    assert not torch.equal(a._xyz, sc["xyz"].to(DEV))


def test_adam_in_backward_at_full_size_with_atomics(c2):
    """Float-atomic reductions (the benchmark's mode): one step each way.
    After one Adam step exp_avg = (1 - beta1) g and exp_avg_sq = (1 - beta2)
    g^2, so the moments carry the two gradients, which agree to the gradient
    tolerance; the first update is lr * g / |g| per entry, so a parameter
    moves by lr either way and the two can differ only where the gradient's
    sign does (by at most 2 lr)."""
    print("C2 adam-in-backward, atomics", flush=True)
    sc = c2["sc"]
    a, oa = _train(sc, 1, False, False)
    b, ob = _train(sc, 1, True, False)
    lrs = {g["params"][0].data_ptr(): g for g in oa.param_groups}
    for n in NAMES:
        pa, pb = getattr(a, n), getattr(b, n)
        sa, sb = oa.state[pa], ob.state[pb]
        assert sa["step"] == sb["step"] == 1, n
        assert_close_grad(n + ".exp_avg", np_(sb["exp_avg"]).ravel(), np_(sa["exp_avg"]).ravel())
        assert_close_grad(n + ".exp_avg_sq", np.sqrt(np_(sb["exp_avg_sq"]).ravel()),
                          np.sqrt(np_(sa["exp_avg_sq"]).ravel()))
        grp = lrs[pa.data_ptr()]
        lr = max(grp["lr"], grp.get("lead_lr", 0.0) or 0.0)
        step = (pb - pa).abs().max().item()
        assert step <= 2.0 * lr * (1 + 1e-5) + 1e-6 * pa.abs().max().item(), (n, step, lr)
        flipped = ((pb - pa).abs() > 1e-6 * pa.abs().max().item()).float().mean().item()
        assert flipped < 1e-3, (n, flipped)
