"""Fused L1 + SSIM loss (HIP) against the reference formula of
trainer/losses.py:98-251 evaluated by torch in float64 on the CPU.

Tolerance: loss and its SSIM / L1 parts to 2e-6 absolute; the gradient to
1e-3 relative of its largest entry (fp32 separable window sums vs fp64 2-D
convolution)."""
import numpy as np
import pytest
import torch

from gs_hip.loss import fused_photometric_loss, photometric_loss, ssim

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _reference(pred, gt):
    p = pred.detach().double().cpu().requires_grad_(True)
    loss = photometric_loss(p, gt.double().cpu())
    loss.backward()
    return float(loss), float(ssim(p.detach(), gt.double().cpu())), p.grad.numpy()


@pytest.mark.parametrize("C,H,W,seed", [(3, 540, 960, 0), (3, 37, 23, 1), (1, 16, 16, 2), (3, 270, 480, 3)])
def test_fused_loss_matches_reference(C, H, W, seed):
    g = torch.Generator().manual_seed(seed)
    gt = torch.rand((C, H, W), generator=g)
    pred = (gt + 0.2 * torch.randn((C, H, W), generator=g)).clamp(0, 1)
    pred[:, : H // 3] = gt[:, : H // 3]  # some exactly equal pixels (sign 0 in the L1 term)
    ref_loss, ref_ssim, ref_grad = _reference(pred, gt)
    x = pred.to(DEV).requires_grad_(True)
    loss, stats = fused_photometric_loss(x, gt.to(DEV))
    loss.backward()
    torch.cuda.synchronize()
    assert abs(float(loss) - ref_loss) < 2e-6, (float(loss), ref_loss)
    assert abs(float(stats[1]) - ref_ssim) < 2e-6
    gx = x.grad.cpu().numpy()
    scale = np.abs(ref_grad).max()
    assert np.abs(gx - ref_grad).max() <= 1e-3 * scale, np.abs(gx - ref_grad).max() / scale


def test_fused_loss_upstream_gradient_scales():
    g = torch.Generator().manual_seed(5)
    gt = torch.rand((3, 64, 80), generator=g)
    pred = torch.rand((3, 64, 80), generator=g)
    x1 = pred.to(DEV).requires_grad_(True)
    l1, _ = fused_photometric_loss(x1, gt.to(DEV))
    (3.5 * l1).backward()
    x2 = pred.to(DEV).requires_grad_(True)
    l2, _ = fused_photometric_loss(x2, gt.to(DEV))
    l2.backward()
    torch.testing.assert_close(x1.grad, 3.5 * x2.grad, rtol=1e-6, atol=0)
