"""Photometric loss of the CF-3DGS training step, in torch.

``Loss.forward`` (trainer/losses.py:98-136) with the default config:
  (1 - lambda_dssim) * L1 + lambda_dssim * (1 - SSIM), lambda_dssim = 0.2
(arguments/__init__.py:113), the depth term off (lambda_depth = 0, :114).
SSIM is ``SSIM_V2`` (trainer/losses.py:212-251): 11x11 Gaussian window,
sigma 1.5 (:143-159), zero padding, C1 = 0.01^2, C2 = 0.03^2 (:164-209).
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

_WINDOWS = {}


def gaussian_window(window_size: int = 11, sigma: float = 1.5, channel: int = 3, device=None,
                    dtype=torch.float32) -> torch.Tensor:
    key = (window_size, sigma, channel, str(device), dtype)
    w = _WINDOWS.get(key)
    if w is None:
        g = torch.tensor([math.exp(-(x - window_size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(window_size)])
        g = g / g.sum()
        w2 = (g[:, None] @ g[None, :]).float()
        w = w2.expand(channel, 1, window_size, window_size).contiguous().to(device=device, dtype=dtype)
        _WINDOWS[key] = w
    return w


def ssim(img1: torch.Tensor, img2: torch.Tensor, window_size: int = 11) -> torch.Tensor:
    if img1.dim() == 3:
        img1 = img1.unsqueeze(0)
    if img2.dim() == 3:
        img2 = img2.unsqueeze(0)
    channel = img1.shape[1]
    w = gaussian_window(window_size, 1.5, channel, img1.device, img1.dtype)
    pad = window_size // 2
    mu1 = F.conv2d(img1, w, padding=pad, groups=channel)
    mu2 = F.conv2d(img2, w, padding=pad, groups=channel)
    mu1_sq, mu2_sq, mu1_mu2 = mu1.pow(2), mu2.pow(2), mu1 * mu2
    s1 = F.conv2d(img1 * img1, w, padding=pad, groups=channel) - mu1_sq
    s2 = F.conv2d(img2 * img2, w, padding=pad, groups=channel) - mu2_sq
    s12 = F.conv2d(img1 * img2, w, padding=pad, groups=channel) - mu1_mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu1_mu2 + C1) * (2 * s12 + C2)) / ((mu1_sq + mu2_sq + C1) * (s1 + s2 + C2))
    return m.mean()


def l1_loss(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    return torch.abs(a - b).mean()


def photometric_loss(pred: torch.Tensor, gt: torch.Tensor, lambda_dssim: float = 0.2) -> torch.Tensor:
    """The reference formula in torch ops (conv2d windows)."""
    return (1.0 - lambda_dssim) * l1_loss(pred, gt) + lambda_dssim * (1.0 - ssim(pred, gt))


def _window_taps(window_size: int = 11, sigma: float = 1.5):
    """The 1-D taps exactly as trainer/losses.py:143-146 forms them (float32)."""
    g = torch.tensor([math.exp(-(x - window_size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(window_size)])
    g = g / g.sum()
    import ctypes
    return (ctypes.c_float * window_size)(*[float(v) for v in g])


_TAPS = None


class _FusedPhotometric(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pred, gt, lambda_dssim, clamp01):
        import ctypes
        from . import _lib
        global _TAPS
        if _TAPS is None:
            _TAPS = _window_taps()
        lib = _lib.load()
        if pred.device.type != "cuda":
            raise RuntimeError("the fused loss runs on the GPU only")
        x = pred.contiguous().float()
        y = gt.to(x.device).contiguous().float()
        if x.dim() == 4:
            x, y = x[0], y[0]
        C, H, W = x.shape
        ws = torch.empty(lib.gs_loss_workspace_bytes(C, H, W), dtype=torch.uint8, device=x.device)
        out = torch.empty(3, dtype=torch.float32, device=x.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
        _lib.check(lib.gs_ssim_l1_forward(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), C, H, W,
                                          _TAPS, float(lambda_dssim), int(bool(clamp01)),
                                          ctypes.c_void_p(out.data_ptr()),
                                          ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream), "gs_ssim_l1_forward")
        ctx.save_for_backward(x, y, ws)
        ctx.shape = pred.shape
        ctx.lam = float(lambda_dssim)
        ctx.clamp01 = int(bool(clamp01))
        ctx.mark_non_differentiable(out)
        ctx.set_materialize_grads(False)
        return out[0], out

    @staticmethod
    def backward(ctx, g_loss, g_out):
        import ctypes
        from . import _lib
        lib = _lib.load()
        if g_loss is None:
            return None, None, None, None
        x, y, ws = ctx.saved_tensors
        C, H, W = x.shape
        go = g_loss.reshape(1).contiguous().float()
        dx = torch.empty_like(x)
        stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
        _lib.check(lib.gs_ssim_l1_backward(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), C, H, W,
                                           _TAPS, ctx.lam, ctx.clamp01, ctypes.c_void_p(go.data_ptr()),
                                           ctypes.c_void_p(ws.data_ptr()), ws.numel(), ctypes.c_void_p(dx.data_ptr()),
                                           stream), "gs_ssim_l1_backward")
        return dx.view(ctx.shape), None, None, None


def fused_photometric_loss(pred: torch.Tensor, gt: torch.Tensor, lambda_dssim: float = 0.2, clamp01: bool = False):
    """(1 - l) L1 + l (1 - SSIM) in two HIP kernels forward and one backward.
    Returns (loss, stats) with stats = [loss, mean SSIM, mean L1] on the device."""
    return _FusedPhotometric.apply(pred, gt, lambda_dssim, clamp01)


def fused_photometric_value_and_grad(pred: torch.Tensor, gt: torch.Tensor, lambda_dssim: float = 0.2,
                                     clamp01: bool = False):
    """Forward and backward of the fused loss in one call, outside autograd:
    returns (stats [loss, SSIM, L1] on the device, d loss / d pred)."""
    import ctypes
    from . import _lib
    global _TAPS
    if _TAPS is None:
        _TAPS = _window_taps()
    lib = _lib.load()
    if pred.device.type != "cuda":
        raise RuntimeError("the fused loss runs on the GPU only")
    x = pred.contiguous().float()
    y = gt.to(x.device).contiguous().float()
    C, H, W = x.shape[-3:]
    ws = torch.empty(lib.gs_loss_workspace_bytes(C, H, W), dtype=torch.uint8, device=x.device)
    out = torch.empty(3, dtype=torch.float32, device=x.device)
    one = torch.ones(1, dtype=torch.float32, device=x.device)
    dx = torch.empty_like(x)
    stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(lib.gs_ssim_l1_forward(p(x), p(y), C, H, W, _TAPS, float(lambda_dssim), int(bool(clamp01)), p(out),
                                      p(ws), ws.numel(), stream), "gs_ssim_l1_forward")
    _lib.check(lib.gs_ssim_l1_backward(p(x), p(y), C, H, W, _TAPS, float(lambda_dssim), int(bool(clamp01)), p(one),
                                       p(ws), ws.numel(), p(dx), stream), "gs_ssim_l1_backward")
    return out, dx.view(pred.shape)
