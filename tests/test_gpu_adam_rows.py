"""The Adam step inside the rasteriser's backward (gs_backward_adam) for every
SH row shape it accepts.

The fused kernel keeps a Gaussian's SH gradient row in registers and moves
parameter rows through LDS.  Its branches depend on the row width 3M (a
multiple of 4 or not: M = 1, 9, 16), on the active degree D (coefficients past
(D+1)^2 get a zero gradient, but their moments still decay, as
torch.optim.Adam does for a zero gradient), and on a partial last block
(P not a multiple of 256).  Each case is compared bit for bit with
gs_backward followed by gs_adam_step_rows (FusedAdam.step) over three steps,
with the fixed-order reductions, so the gradients are the same on both sides.
The SH learning rate has the f_dc lead (3 floats at feature_lr, the rest at
feature_lr / 20), as training_setup sets it.
"""
import pytest
import torch

from gs_hip import camera as gcam
from gs_hip import raster, scenes
from gs_hip.optim import FusedAdam

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
P_RAGGED = 3001


LR_XYZ, LR_DC, LR_REST = 1.6e-4, 2.5e-3, 2.5e-3 / 20.0


def _steps(M, fused, degrees):
    """Steps of the means and SH coefficients over ``degrees`` (the active
    degree of each step).  Fused: one [P, M, 3] tensor stepped inside the
    backward, f_dc at LR_DC and the rest at LR_REST.  Separate: gs_backward's
    gradients, then FusedAdam.step over f_dc [P, 1, 3] and f_rest [P, M-1, 3]
    as two tensors (the reference's two groups; Adam is elementwise, so the
    split does not change any element's update).  Returns (means, shs, the
    SH moments, the means' moments, step counts)."""
    sc = scenes.plain_scene(seed=100 + M, P=P_RAGGED, W=160, H=112)
    means = sc["means"].to(DEV).clone()
    sh0 = sc["shs"][:, :M].contiguous().to(DEV)
    op, scl, rot = (sc[k].to(DEV) for k in ("opacities", "scales", "rotations"))
    cam, bg = sc["cam"].to(DEV), sc["bg"].to(DEV)
    g = torch.Generator().manual_seed(M)
    if fused:
        shs = sh0.clone()
        sh_group = {"params": [shs], "lr": LR_DC} if M == 1 else \
            {"params": [shs], "lr": LR_REST, "lead_lr": LR_DC, "lead": 3, "row": 3 * M}
        opt = FusedAdam([{"params": [means], "lr": LR_XYZ}, sh_group], eps=1e-15)
    else:
        dc, rest = sh0[:, :1].clone(), sh0[:, 1:].clone()
        groups = [{"params": [means], "lr": LR_XYZ}, {"params": [dc], "lr": LR_DC}]
        if M > 1:
            groups.append({"params": [rest], "lr": LR_REST})
        opt = FusedAdam(groups, eps=1e-15)
    for d in degrees:
        rs = gcam.raster_settings(cam, bg, d)
        cur = shs if fused else torch.cat([dc, rest], 1).contiguous()
        args = (means, cur, None, op, scl, rot, None)
        _, _, _, _, st = raster.forward(rs, *args)
        dcot = ((torch.rand((3, 112, 160), generator=g) - 0.5) / (112 * 160)).to(DEV)
        if fused:
            raster.backward(rs, st, *args, None, dcot, deterministic=True, adam=opt.in_backward(means, shs).take())
        else:
            gr = raster.backward(rs, st, *args, None, dcot, deterministic=True)
            means.grad = gr["means3D"]
            dc.grad = gr["shs"][:, :1].contiguous()
            rest.grad = gr["shs"][:, 1:].contiguous() if M > 1 else None
            opt.step()
            means.grad = dc.grad = rest.grad = None
    torch.cuda.synchronize()
    st_m = opt.state[means]
    if fused:
        st_s = opt.state[shs]
        sh_m, sh_v, sh_steps = st_s["exp_avg"], st_s["exp_avg_sq"], [st_s["step"]]
    else:
        parts = [opt.state[dc]] + ([opt.state[rest]] if M > 1 else [])
        sh_m = torch.cat([q["exp_avg"] for q in parts], 1)
        sh_v = torch.cat([q["exp_avg_sq"] for q in parts], 1)
        sh_steps = [q["step"] for q in parts]
        shs = torch.cat([dc, rest], 1)
    return means, shs, (sh_m, sh_v), (st_m["exp_avg"], st_m["exp_avg_sq"]), [st_m["step"]] + sh_steps


@pytest.mark.parametrize("M,degrees", [
    (1, (0, 0, 0)),
    (9, (0, 1, 2)),
    (9, (2, 1, 1)),      # degree drops: the dropped coefficients' moments decay
    (16, (0, 0, 0)),
    (16, (1, 1, 1)),
    (16, (3, 3, 3)),
    (16, (0, 1, 3)),
])
def test_adam_in_backward_rows_equal_the_separate_step(M, degrees):
    print("adam in backward: M", M, "degrees", degrees, flush=True)
    a_means, a_shs, a_sh_mom, a_mom, a_steps = _steps(M, False, degrees)
    b_means, b_shs, b_sh_mom, b_mom, b_steps = _steps(M, True, degrees)
    assert set(a_steps) == set(b_steps) == {len(degrees)}
    assert torch.equal(a_means, b_means)
    assert torch.equal(a_shs, b_shs)
    for x, y in zip(a_sh_mom + a_mom, b_sh_mom + b_mom):
        assert torch.equal(x, y)
    # the steps moved the coefficients that the active degrees read
    nc = (max(degrees) + 1) ** 2
    sh0 = scenes.plain_scene(seed=100 + M, P=P_RAGGED, W=160, H=112)["shs"][:, :nc].to(DEV)
    assert not torch.equal(a_shs[:, :nc], sh0)
