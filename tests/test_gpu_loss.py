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


@pytest.mark.parametrize("C,H,W,clamp", [(3, 540, 960, 1), (3, 37, 23, 0)])
def test_forward_backward_in_one_call_equals_the_two_calls(C, H, W, clamp):
    """gs_ssim_l1_forward_backward (the backward's first workgroup reduces the
    forward's partials) gives the loss statistics and the gradient of
    gs_ssim_l1_forward + gs_ssim_l1_backward bit for bit."""
    import ctypes
    from gs_hip import _lib
    from gs_hip.loss import _window_taps
    lib = _lib.load()
    g = torch.Generator().manual_seed(7)
    y = torch.rand((C, H, W), generator=g).to(DEV)
    x = (torch.rand((C, H, W), generator=g) * 1.4 - 0.2).to(DEV)
    one = torch.full((1,), 0.75, device=DEV)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    nb = lib.gs_loss_workspace_bytes(C, H, W)
    taps = _window_taps()
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    ws1, ws2 = (torch.empty(nb, dtype=torch.uint8, device=DEV) for _ in range(2))
    out1, out2 = torch.zeros(3, device=DEV), torch.zeros(3, device=DEV)
    dx1, dx2 = torch.zeros_like(x), torch.zeros_like(x)
    assert lib.gs_ssim_l1_forward(p(x), p(y), C, H, W, taps, 0.2, clamp, p(out1), p(ws1), nb, stream) == 0
    assert lib.gs_ssim_l1_backward(p(x), p(y), C, H, W, taps, 0.2, clamp, p(one), p(ws1), nb, p(dx1), stream) == 0
    assert lib.gs_ssim_l1_forward_backward(p(x), p(y), C, H, W, taps, 0.2, clamp, p(out2), p(one), p(ws2), nb,
                                           p(dx2), stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(out1, out2) and torch.equal(dx1, dx2)
    assert float(dx1.abs().max()) > 0
    assert lib.gs_ssim_l1_forward_backward(p(x), p(y), C, H, W, taps, 0.2, clamp, p(out2), p(one), p(ws2), nb - 1,
                                           p(dx2), stream) == 2  # GS_E_WORKSPACE


@pytest.mark.parametrize("clamp", [0, 1])
def test_multi_image_loss_equals_each_images_own_call(clamp):
    """gs_multi_loss (every image's forward and backward blocks in one launch
    of each kernel) gives each image the statistics and the gradient of its
    own gs_ssim_l1_forward_backward call, bit for bit; the images differ in
    size (tiles cut at the right and bottom edges) and channel count."""
    import ctypes
    from gs_hip import _lib
    from gs_hip.loss import _window_taps
    lib = _lib.load()
    shapes = [(3, 90, 160), (3, 270, 480), (1, 37, 53), (3, 540, 960)]
    n = len(shapes)
    g = torch.Generator().manual_seed(11)
    xs = [(torch.rand(s, generator=g) * 1.4 - 0.2).to(DEV) for s in shapes]
    ys = [torch.rand(s, generator=g).to(DEV) for s in shapes]
    one = torch.full((1,), 0.75, device=DEV)
    taps = _window_taps()
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    nbs = [lib.gs_loss_workspace_bytes(*s) for s in shapes]
    ws1 = [torch.empty(b, dtype=torch.uint8, device=DEV) for b in nbs]
    ws2 = [torch.empty(b, dtype=torch.uint8, device=DEV) for b in nbs]
    out1, out2 = torch.zeros((n, 3), device=DEV), torch.zeros((n, 3), device=DEV)
    dx1 = [torch.zeros_like(x) for x in xs]
    dx2 = [torch.zeros_like(x) for x in xs]
    for k, (C, H, W) in enumerate(shapes):
        assert lib.gs_ssim_l1_forward_backward(p(xs[k]), p(ys[k]), C, H, W, taps, 0.2, clamp, p(out1[k]), p(one),
                                               p(ws1[k]), nbs[k], p(dx1[k]), stream) == 0
    vp = ctypes.c_void_p
    arr = lambda ts: (vp * n)(*[t.data_ptr() for t in ts])  # noqa: E731
    ints = lambda v: (ctypes.c_int * n)(*v)  # noqa: E731
    sb = lib.gs_multi_loss_bytes(n)
    host = (ctypes.c_uint8 * sb)()
    blocks = ctypes.c_uint(0)
    assert lib.gs_multi_loss_build(n, arr(xs), arr(ys), ints([s[0] for s in shapes]), ints([s[1] for s in shapes]),
                                   ints([s[2] for s in shapes]), taps, 0.2, clamp, arr([out2[k] for k in range(n)]),
                                   p(one), arr(ws2), (ctypes.c_size_t * n)(*nbs), arr(dx2), host, sb,
                                   ctypes.byref(blocks)) == 0
    dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(DEV)
    assert lib.gs_multi_loss(n, p(dev), blocks.value, stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(out1, out2)
    for k in range(n):
        assert torch.equal(dx1[k], dx2[k]), k
    assert float(out1[:, 0].abs().min()) > 0
