"""Gaussian parameters and the render call, as CF-3DGS's model layer uses the
rasteriser.

  * raw parameters and activations: scene/gaussian_model_cf.py:41-56,
    115-172 (exp scales, normalised rotations, sigmoid opacities,
    cat(features_dc, features_rest) -> [P, 16, 3]);
  * optimiser groups: ``training_setup`` (:242-268) with the defaults of
    arguments/__init__.py:103-111, Adam eps 1e-15;
  * ``render``: CF3DGS_Render.render (:777-910) -- zero screen-space means
    with retained grad, keyword call of the rasteriser, clamp of the image.
The rigid pose of the pose-solve loop is passed to the rasteriser (fused in
its preprocess) instead of transforming the means in torch.
"""
from __future__ import annotations

import torch
import torch.nn as nn
import torch.nn.functional as F

from diff_gaussian_rasterization import GaussianRasterizer

from .camera import Camera, raster_settings


class GaussianParams:
    def __init__(self, xyz, features_dc, features_rest, scaling, rotation, opacity, sh_degree: int,
                 device="cuda"):
        def p(t):
            return nn.Parameter(t.to(device).contiguous().requires_grad_(True))
        self._xyz = p(xyz)
        self._features_dc = p(features_dc)
        self._features_rest = p(features_rest)
        self._scaling = p(scaling)
        self._rotation = p(rotation)
        self._opacity = p(opacity)
        self.active_sh_degree = sh_degree
        self.optimizer = None

    @property
    def get_xyz(self):
        return self._xyz

    @property
    def get_scaling(self):
        return torch.exp(self._scaling)

    @property
    def get_rotation(self):
        return F.normalize(self._rotation)

    @property
    def get_opacity(self):
        return torch.sigmoid(self._opacity)

    @property
    def get_features(self):
        return torch.cat((self._features_dc, self._features_rest), dim=1)

    def training_setup(self, spatial_lr_scale: float = 1.0, position_lr_init: float = 0.00016,
                       feature_lr: float = 0.0025, opacity_lr: float = 0.05, scaling_lr: float = 0.005,
                       rotation_lr: float = 0.001, fused: bool = True):
        groups = [
            {"params": [self._xyz], "lr": position_lr_init * spatial_lr_scale, "name": "xyz"},
            {"params": [self._features_dc], "lr": feature_lr, "name": "f_dc"},
            {"params": [self._features_rest], "lr": feature_lr / 20.0, "name": "f_rest"},
            {"params": [self._opacity], "lr": opacity_lr, "name": "opacity"},
            {"params": [self._scaling], "lr": scaling_lr, "name": "scaling"},
            {"params": [self._rotation], "lr": rotation_lr, "name": "rotation"},
        ]
        if fused:
            from .optim import FusedAdam
            self.optimizer = FusedAdam(groups, lr=0.0, eps=1e-15)
        else:
            self.optimizer = torch.optim.Adam(groups, lr=0.0, eps=1e-15)
        return self.optimizer


def render(gaussians: GaussianParams, cam: Camera, bg: torch.Tensor, pose=None, scaling_modifier: float = 1.0):
    """CF3DGS_Render.render with the rasteriser's 3-tuple return."""
    means3D = gaussians.get_xyz
    screenspace = torch.zeros_like(means3D, requires_grad=True) + 0
    try:
        screenspace.retain_grad()
    except RuntimeError:
        pass
    rs = raster_settings(cam, bg, gaussians.active_sh_degree, scaling_modifier)
    rasterizer = GaussianRasterizer(raster_settings=rs)
    color, radii, depth = rasterizer(means3D=means3D, means2D=screenspace, shs=gaussians.get_features,
                                     colors_precomp=None, opacities=gaussians.get_opacity,
                                     scales=gaussians.get_scaling, rotations=gaussians.get_rotation,
                                     cov3D_precomp=None, pose=pose)
    return {"image": color.clamp(0, 1), "depth": depth, "viewspace_points": screenspace,
            "visibility_filter": radii > 0, "radii": radii}
