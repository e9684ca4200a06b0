"""Drop-in `diff_gaussian_rasterization` for MI355X.

Same Python surface as the package CF-3DGS imports
(scene/gaussian_model_cf.py:28-31, gaussian_renderer/__init__.py:12):

  * ``GaussianRasterizationSettings`` -- NamedTuple with the fields built by
    keyword at scene/gaussian_model_cf.py:811-824;
  * ``GaussianRasterizer(raster_settings)`` -- ``nn.Module`` whose ``forward``
    takes ``means3D, means2D, opacities, shs, colors_precomp, scales,
    rotations, cov3D_precomp`` by keyword (:873-882) and returns
    ``(color [3,H,W], radii [P] int32, depth [1,H,W])``; the caller's 3-tuple
    branch (:897-910).  ``return_alpha=True`` adds the alpha image as a 4th
    output (the 4-tuple branch, :883-896);
  * ``markVisible(positions)``;
  * ``rasterize_gaussians(...)`` and the autograd function.

Argument errors raise ``Exception`` with the upstream wording.  Gradients flow
to means3D, means2D (``.grad`` gets the NDC screen-space gradient used by
densification, :716-719), shs / colors_precomp, opacities, scales /
rotations / cov3D_precomp.  The optional ``pose`` keyword ([tx, ty, tz, qx,
qy, qz, qw]) applies the rigid transform that CF-3DGS performs with
``P[0].retr().act(xyz)`` (:123-133) inside the kernel and returns its
gradient.  The optional ``adam_in_backward`` keyword
(``gs_hip.optim.FusedAdam.in_backward(xyz, shs)``) makes the backward apply
the optimiser's Adam step to means3D and shs itself, in place, instead of
returning their gradients (the step after it then skips them).

All work runs in libgs_hip.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

from typing import NamedTuple

import torch
import torch.nn as nn

try:
    from gs_hip import raster as _raster
except ImportError:  # imported as a sibling package of gs_hip
    import os as _os
    import sys as _sys
    _sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
    from gs_hip import raster as _raster

__all__ = ["GaussianRasterizationSettings", "GaussianRasterizer", "rasterize_gaussians"]


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    prefiltered: bool
    debug: bool


def _none_if_empty(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor) and t.numel() == 0:
        return None
    return t


class _RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp, pose,
                raster_settings, return_alpha, adam_in_backward=None):
        sh_ = _none_if_empty(sh)
        col_ = _none_if_empty(colors_precomp)
        sc_ = _none_if_empty(scales)
        ro_ = _none_if_empty(rotations)
        cv_ = _none_if_empty(cov3Ds_precomp)
        po_ = _none_if_empty(pose)
        color, radii, depth, alpha, st = _raster.forward(
            raster_settings, means3D, sh_, col_, opacities, sc_, ro_, cv_, pose=po_, want_alpha=True)
        if raster_settings.debug:
            torch.cuda.synchronize(means3D.device)
        ctx.raster_settings = raster_settings
        ctx.state = st
        ctx.adam = adam_in_backward
        ctx.has = (sh_ is not None, col_ is not None, sc_ is not None, ro_ is not None, cv_ is not None,
                   po_ is not None)
        ctx.save_for_backward(means3D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp, pose)
        ctx.mark_non_differentiable(radii)
        ctx.set_materialize_grads(False)  # an unused output's gradient stays None (treated as zero)
        if return_alpha:
            return color, radii, depth, alpha
        return color, radii, depth

    @staticmethod
    def backward(ctx, grad_color, grad_radii, grad_depth, *rest):
        grad_alpha = rest[0] if rest else None
        means3D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp, pose = ctx.saved_tensors
        has_sh, has_col, has_sc, has_ro, has_cv, has_pose = ctx.has
        st = ctx.state
        # Only the gradients autograd asks for are written (a pose-only solve
        # skips every per-Gaussian output).
        need = ctx.needs_input_grad
        names = ("means3D", "means2D", "shs", "colors", "opacities", "scales", "rotations", "cov3D", "pose")
        want = {n for n, k in zip(names, need[:9]) if k}
        # adam_in_backward: the means and SH coefficients take their optimiser
        # step here and pass no gradient on
        adam = ctx.adam.take() if ctx.adam is not None and grad_color is not None else None
        g = _raster.backward(
            ctx.raster_settings, st, means3D,
            sh if has_sh else None, colors_precomp if has_col else None, opacities,
            scales if has_sc else None, rotations if has_ro else None, cov3Ds_precomp if has_cv else None,
            pose if has_pose else None, grad_color, grad_depth, grad_alpha, want=want, adam=adam)
        if ctx.raster_settings.debug:
            torch.cuda.synchronize(means3D.device)

        def shaped(name, like):
            t = g[name]
            return t.view_as(like) if t is not None else None

        grad_means3D = shaped("means3D", means3D)
        grad_means2D = g["means2D"]
        grad_sh = shaped("shs", sh) if has_sh else None
        grad_col = shaped("colors", colors_precomp) if has_col else None
        grad_opac = shaped("opacities", opacities)
        grad_sc = shaped("scales", scales) if has_sc else None
        grad_ro = shaped("rotations", rotations) if has_ro else None
        grad_cv = shaped("cov3D", cov3Ds_precomp) if has_cv else None
        grad_pose = shaped("pose", pose) if has_pose else None
        return (grad_means3D, grad_means2D, grad_sh, grad_col, grad_opac, grad_sc, grad_ro, grad_cv, grad_pose,
                None, None, None)


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                        raster_settings, pose=None, return_alpha=False, adam_in_backward=None):
    empty = torch.empty(0, device=means3D.device)
    return _RasterizeGaussians.apply(
        means3D, means2D,
        sh if sh is not None else empty,
        colors_precomp if colors_precomp is not None else empty,
        opacities,
        scales if scales is not None else empty,
        rotations if rotations is not None else empty,
        cov3Ds_precomp if cov3Ds_precomp is not None else empty,
        pose if pose is not None else empty,
        raster_settings, bool(return_alpha), adam_in_backward)


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings, return_alpha: bool = False):
        super().__init__()
        self.raster_settings = raster_settings
        self.return_alpha = return_alpha

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            return _raster.mark_visible(positions, rs.viewmatrix, rs.projmatrix)

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None, rotations=None,
                cov3D_precomp=None, pose=None, adam_in_backward=None):
        rs = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception('Please provide excatly one of either SHs or precomputed colors!')
        if ((scales is None or rotations is None) and cov3D_precomp is None) or \
                ((scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception('Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!')
        if means3D.dim() != 2 or means3D.shape[1] != 3:
            raise Exception('means3D must have dimensions (num_points, 3)')
        if adam_in_backward is not None and (pose is not None or shs is None):
            raise Exception('adam_in_backward needs SH colours and no pose')
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations,
                                   cov3D_precomp, rs, pose=pose, return_alpha=self.return_alpha,
                                   adam_in_backward=adam_in_backward)
