"""Adaptive density control of ``GaussianParams`` on the device.

The reference's CFGaussianModel (scene/gaussian_model_cf.py:523-719) and its
trainer (trainer/cf3dgs_trainer.py:144-160) keep, per Gaussian, the summed
screen-space gradient norm, a visibility count and the largest screen radius,
and every ``densification_interval`` steps

  * clone the Gaussians with a large mean gradient and a small scale,
  * split those with a large mean gradient and a large scale into two
    samples of their own distribution (scale / 1.6), dropping the original,
  * prune the transparent (opacity < 0.005), and, after the first opacity
    reset, the too large,

re-indexing every parameter and its Adam moments each time.  Here the
statistics are one kernel per step (``gs_densify_stats``) and the three
re-indexings are one stable compaction (``gs_densify_mark`` + ``gs_densify_emit``)
that writes every surviving row of every parameter and moment once, in the
order the reference's sequence of ``torch.cat`` and boolean indexing leaves
it.  The split samples are drawn with the reference's own
``torch.normal(mean=zeros, std=stds)`` call, so a caller's generator state
gives the same points.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch
import torch.nn as nn

from . import _lib
from . import raster as _raster

PARAMS = ("_xyz", "_features", "_opacity", "_scaling", "_rotation")


def _p(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def init_stats(g) -> None:
    """Zero statistics for the current Gaussian count (training_setup,
    scene/gaussian_model_cf.py:244-247, and densification_postfix)."""
    P = g._xyz.shape[0]
    dev = g._xyz.device
    g.xyz_gradient_accum = torch.zeros((P, 1), dtype=torch.float32, device=dev)
    g.denom = torch.zeros((P, 1), dtype=torch.float32, device=dev)
    g.max_radii2D = torch.zeros(P, dtype=torch.float32, device=dev)


def add_densification_stats(g, viewspace_points: torch.Tensor, radii: torch.Tensor) -> None:
    """For every Gaussian with radius > 0 (the visibility filter): grad_accum
    += |dL/dmean2D.xy|, denom += 1, max_radii2D = max(max_radii2D, radius)
    (trainer/cf3dgs_trainer.py:144-152, gaussian_model_cf.py:716-719)."""
    grad = viewspace_points.grad if isinstance(viewspace_points, torch.Tensor) and \
        viewspace_points.grad is not None else viewspace_points
    if grad is None:
        raise RuntimeError("add_densification_stats: the screen-space points have no gradient")
    _raster.require_device(grad, "viewspace_points")
    P = g._xyz.shape[0]
    if grad.shape[0] != P or radii.shape[0] != P:
        raise ValueError("add_densification_stats: sizes differ from the Gaussian count")
    if getattr(g, "denom", None) is None or g.denom.shape[0] != P:
        init_stats(g)
    grad = grad.detach().contiguous().float()
    radii = radii.detach().to(torch.int32).contiguous()
    lib = _lib.load()
    _lib.check(lib.gs_densify_stats(P, _p(grad), _p(radii), _p(g.xyz_gradient_accum), _p(g.denom),
                                    _p(g.max_radii2D), _raster._stream(grad.device)), "gs_densify_stats")


def _apply(g, densify: bool, prune: bool, max_grad: float, min_opacity: float, extent: float,
           max_screen: float, big_extent: float, generator: Optional[torch.Generator]) -> dict:
    lib = _lib.load()
    dev = g._xyz.device
    P = g._xyz.shape[0]
    M = g._features.shape[1]
    stream = _raster._stream(dev)
    if densify and (getattr(g, "denom", None) is None or g.denom.shape[0] != P):
        init_stats(g)
    flags = torch.empty(max(P, 4) + 4, dtype=torch.uint8, device=dev)
    totals = torch.zeros(4, dtype=torch.int32, device=dev)
    ws = torch.empty(max(int(lib.gs_densify_workspace_bytes(P)), 256), dtype=torch.uint8, device=dev)
    scaling = g._scaling.detach().contiguous()
    opacity = g._opacity.detach().contiguous()
    # thresholds as torch compares a float32 tensor with a Python float
    f32 = lambda x: float(torch.tensor(x, dtype=torch.float32))  # noqa: E731
    _lib.check(lib.gs_densify_mark(
        P, _p(g.xyz_gradient_accum if densify else None), _p(g.denom if densify else None), _p(scaling),
        _p(opacity), _p(getattr(g, "max_radii2D", None) if max_screen > 0 else None), f32(max_grad),
        f32(g.percent_dense * extent), f32(min_opacity), f32(big_extent), f32(max_screen), int(densify), int(prune),
        _p(flags), _p(totals), _p(ws), ws.numel(), stream), "gs_densify_mark")
    K_o, K_c, K_s, K_sel = (int(v) for v in totals.cpu())
    samples = None
    if K_sel:
        # the reference's draw (densify_and_split, N = 2): std = get_scaling of
        # the selected Gaussians, repeated once per copy
        sel = (flags[:P] & 4) != 0
        stds = torch.exp(g._scaling.detach()[sel]).repeat(2, 1)
        means = torch.zeros((stds.size(0), 3), device=dev)
        samples = torch.normal(mean=means, std=stds, generator=generator).contiguous()
    P_new = K_o + K_c + 2 * K_s
    opt = g.optimizer
    params = [getattr(g, n) for n in PARAMS]
    states = [opt.state.get(p) if opt is not None else None for p in params]
    moments = all(st is not None and "exp_avg" in st for st in states)
    srcs = [p.detach().contiguous() for p in params]
    dsts = [torch.empty((P_new,) + tuple(p.shape[1:]), dtype=torch.float32, device=dev) for p in params]
    src_m = [st["exp_avg"] for st in states] if moments else [None] * 5
    src_v = [st["exp_avg_sq"] for st in states] if moments else [None] * 5
    dst_m = [torch.empty_like(d) for d in dsts] if moments else [None] * 5
    dst_v = [torch.empty_like(d) for d in dsts] if moments else [None] * 5
    src_arr = (ctypes.c_void_p * 15)(*[t.data_ptr() if t is not None else None for t in srcs + src_m + src_v])
    dst_arr = (ctypes.c_void_p * 15)(*[t.data_ptr() if t is not None else None for t in dsts + dst_m + dst_v])
    _lib.check(lib.gs_densify_emit(P, M, src_arr, dst_arr, _p(samples), _p(flags), _p(totals), _p(ws), ws.numel(),
                                   stream), "gs_densify_emit")
    # swap the new tensors into the model and the optimiser (same groups, same
    # step counts; the moments are the compacted ones)
    for name, old, new, st, m, v in zip(PARAMS, params, dsts, states, dst_m, dst_v):
        p_new = nn.Parameter(new.requires_grad_(True))
        if opt is not None:
            for group in opt.param_groups:
                group["params"] = [p_new if q is old else q for q in group["params"]]
            if st is not None:
                del opt.state[old]
                if moments:
                    st["exp_avg"], st["exp_avg_sq"] = m, v
                opt.state[p_new] = st
        setattr(g, name, p_new)
    if opt is not None and hasattr(opt, "_fused_cache"):
        opt._fused_cache.clear()
    if densify:
        init_stats(g)  # densification_postfix resets the statistics
    else:
        keep = (flags[:P] & 1) != 0
        g.xyz_gradient_accum = g.xyz_gradient_accum[keep] if getattr(g, "denom", None) is not None else None
        g.denom = g.denom[keep] if getattr(g, "denom", None) is not None else None
        g.max_radii2D = g.max_radii2D[keep] if getattr(g, "max_radii2D", None) is not None else None
    return dict(kept=K_o, cloned=K_c, split=K_sel, split_kept=K_s, total=P_new)


def densify_and_prune(g, max_grad: float, min_opacity: float, extent: float, max_screen_size=None,
                      generator: Optional[torch.Generator] = None) -> dict:
    """densify_and_prune (scene/gaussian_model_cf.py:681-697).  Its radius
    test reads max_radii2D after densification_postfix has zeroed it, so
    only the opacity and world-size tests can prune here; the compaction
    reproduces that."""
    return _apply(g, True, True, max_grad, min_opacity, extent, 0.0,
                  0.1 * extent if max_screen_size else 0.0, generator)


def densify(g, max_grad: float, min_opacity: float, extent: float, max_screen_size=None,
            generator: Optional[torch.Generator] = None) -> dict:
    """densify (:712-717): clone and split, no prune."""
    return _apply(g, True, False, max_grad, min_opacity, extent, 0.0, 0.0, generator)


def prune(g, max_grad: float, min_opacity: float, extent: float, max_screen_size=None) -> dict:
    """prune (:699-710): opacity, and with max_screen_size the screen radius
    and the world size."""
    return _apply(g, False, True, max_grad, min_opacity, extent, float(max_screen_size or 0.0),
                  0.1 * extent if max_screen_size else 0.0, None)


def reset_opacity(g) -> None:
    """reset_opacity (:452-457): opacity = inverse_sigmoid(min(sigmoid, 0.01))
    with fresh (zero) moments, in torch's arithmetic."""
    with torch.no_grad():
        op = torch.sigmoid(g._opacity)
        x = torch.min(op, torch.ones_like(op) * 0.01)
        new = torch.log(x / (1 - x)).contiguous()
    old = g._opacity
    p_new = nn.Parameter(new.requires_grad_(True))
    opt = g.optimizer
    if opt is not None:
        for group in opt.param_groups:
            group["params"] = [p_new if q is old else q for q in group["params"]]
        st = opt.state.pop(old, None)
        if st is not None:
            st["exp_avg"] = torch.zeros_like(new)
            st["exp_avg_sq"] = torch.zeros_like(new)
            opt.state[p_new] = st
        if hasattr(opt, "_fused_cache"):
            opt._fused_cache.clear()
    g._opacity = p_new
