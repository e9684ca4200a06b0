"""Gaussian parameters and the render call, as CF-3DGS's model layer uses the
rasteriser.

  * raw parameters and activations: scene/gaussian_model_cf.py:48-56,
    116-169 (exp scales, normalised rotations, sigmoid opacities,
    cat(features_dc, features_rest) -> [P, 16, 3]);
  * optimiser groups: ``training_setup`` (:242-268) with the defaults of
    arguments/__init__.py:103-111, Adam eps 1e-15;
  * ``render``: CF3DGS_Render.render (:777-910) -- zero screen-space means
    with retained grad, keyword call of the rasteriser, clamp of the image.
The rigid pose of the pose-solve loop is passed to the rasteriser (fused in
its preprocess) instead of transforming the means in torch.

Two layout choices differ from the reference's module, not its arithmetic:
  * the SH coefficients live in ONE [P, 16, 3] parameter.  The reference
    keeps f_dc [P,1,3] and f_rest [P,15,3] apart and concatenates them on
    every ``get_features`` (a full copy of the SH forward and a split copy of
    its gradient backward); ``_features_dc`` / ``_features_rest`` are views
    here, and the fused Adam gives the leading 3 floats of every 48 the f_dc
    learning rate and the rest the f_rest one;
  * ``render`` evaluates the three activations in one HIP kernel each way
    (``activate``) instead of ~17 torch kernels.  The single-tensor getters
    stay torch expressions for callers outside the step.
"""
from __future__ import annotations

import ctypes

import torch
import torch.nn as nn
import torch.nn.functional as F

from diff_gaussian_rasterization import GaussianRasterizer

from . import _lib
from .camera import Camera, raster_settings


def _p(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _aligned(t: torch.Tensor) -> torch.Tensor:
    t = t.contiguous()
    return t if t.data_ptr() % 16 == 0 else t.clone()


class _Activate(torch.autograd.Function):
    """(exp(raw_scale), normalize(raw_rot), sigmoid(raw_opacity)) in one
    launch forward (gs_activate_forward) and one backward
    (gs_activate_backward)."""

    @staticmethod
    def forward(ctx, raw_scale, raw_rot, raw_opac, adam=None):
        dev = raw_scale.device
        if dev.type != "cuda":
            raise RuntimeError("the fused activations run on the GPU only")
        for name, t in (("raw_scale", raw_scale), ("raw_rot", raw_rot), ("raw_opacity", raw_opac)):
            if t.dtype != torch.float32:
                raise RuntimeError(f"{name} must be float32")
        rs, rr, ro = raw_scale.contiguous(), _aligned(raw_rot), raw_opac.contiguous()
        P = rs.shape[0]
        scales = torch.empty_like(rs)
        rots = torch.empty_like(rr)
        opac = torch.empty_like(ro)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(_lib.load().gs_activate_forward(P, _p(rs), _p(rr), _p(ro), _p(scales), _p(rots), _p(opac),
                                                   stream), "gs_activate_forward")
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(rr, scales, opac)
        # adam in backward updates the caller's raw tensors, not converted copies
        ctx.adam = adam if adam is not None and rs is raw_scale and rr is raw_rot and ro is raw_opac else None
        return scales, rots, opac

    @staticmethod
    def backward(ctx, g_s, g_r, g_o):
        rr, scales, opac = ctx.saved_tensors
        P = rr.shape[0]
        # an output that took no part in the loss passes no gradient on (as in torch)
        need = ctx.needs_input_grad
        stream = ctypes.c_void_p(torch.cuda.current_stream(rr.device).cuda_stream)
        if ctx.adam is not None and all(need[:3]) and g_s is not None and g_r is not None and g_o is not None:
            # one Adam step of the raw parameters here; they get no gradient
            ad = ctx.adam.take()
            rs_p, rr_p, ro_p = ctx.adam.params
            _lib.check(_lib.load().gs_activate_backward_adam(
                P, _p(rs_p), _p(rr_p), _p(ro_p), _p(scales), _p(opac), _p(g_s.contiguous()), _p(_aligned(g_r)),
                _p(g_o.contiguous()), ctypes.byref(ad), stream), "gs_activate_backward_adam")
            return None, None, None, None
        d_s = torch.empty_like(scales) if need[0] and g_s is not None else None
        d_r = torch.empty_like(rr) if need[1] and g_r is not None else None
        d_o = torch.empty_like(opac) if need[2] and g_o is not None else None
        g_s = g_s.contiguous() if g_s is not None else None
        g_r = _aligned(g_r) if g_r is not None else None
        g_o = g_o.contiguous() if g_o is not None else None
        _lib.check(_lib.load().gs_activate_backward(P, _p(rr), _p(scales), _p(opac), _p(g_s), _p(g_r), _p(g_o),
                                                    _p(d_s), _p(d_r), _p(d_o), stream), "gs_activate_backward")
        return d_s, d_r, d_o, None


def activate(raw_scale: torch.Tensor, raw_rot: torch.Tensor, raw_opacity: torch.Tensor, adam=None):
    """The reference's scaling / rotation / opacity activations
    (torch.exp, F.normalize, torch.sigmoid) fused: returns
    (scales [P,3], rotations [P,4], opacities [P,1]).  With ``adam``
    (``FusedAdam.in_backward_activations(raw_scale, raw_rot, raw_opacity)``)
    the backward applies their Adam step in place instead of returning the
    raw gradients."""
    return _Activate.apply(raw_scale, raw_rot, raw_opacity, adam)


class GaussianParams:
    def __init__(self, xyz, features_dc, features_rest, scaling, rotation, opacity, sh_degree: int,
                 device="cuda"):
        def p(t):
            return nn.Parameter(t.to(device).contiguous().requires_grad_(True))
        self._xyz = p(xyz)
        self._features = p(torch.cat((features_dc, features_rest), dim=1))
        self._scaling = p(scaling)
        self._rotation = p(rotation)
        self._opacity = p(opacity)
        self.active_sh_degree = sh_degree
        self.optimizer = None
        self._screen_zeros = None
        self.percent_dense = 0.0
        self.xyz_gradient_accum = self.denom = self.max_radii2D = None

    @property
    def _features_dc(self):
        return self._features[:, :1]

    @property
    def _features_rest(self):
        return self._features[:, 1:]

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
        return self._features

    def screenspace_leaf(self) -> torch.Tensor:
        """A fresh zero [P, 3] leaf for the rasteriser's means2D input, whose
        .grad receives dL/dmean2D.  The reference makes it as
        ``zeros_like(xyz, requires_grad=True) + 0`` plus ``retain_grad()``
        (scene/gaussian_model_cf.py:793-805).  That costs a fill, an add and
        a clone of the gradient every step.  Here every step gets a new leaf
        over one zero buffer, which is never written: the rasteriser only
        routes the gradient through it."""
        xyz = self._xyz
        z = self._screen_zeros
        if z is None or z.shape != xyz.shape or z.device != xyz.device:
            z = self._screen_zeros = torch.zeros_like(xyz)
        return z.detach().requires_grad_(True)

    def get_activated(self, adam=None):
        """(scales, rotations, opacities) in one fused launch."""
        return activate(self._scaling, self._rotation, self._opacity, adam)

    def training_setup(self, spatial_lr_scale: float = 1.0, position_lr_init: float = 0.00016,
                       feature_lr: float = 0.0025, opacity_lr: float = 0.05, scaling_lr: float = 0.005,
                       rotation_lr: float = 0.001, percent_dense: float = 0.01):
        from . import densify
        from .optim import FusedAdam
        self.percent_dense = percent_dense
        densify.init_stats(self)
        row = self._features.shape[1] * self._features.shape[2]
        groups = [
            {"params": [self._xyz], "lr": position_lr_init * spatial_lr_scale, "name": "xyz"},
            # f_rest at feature_lr / 20, its leading 3 floats per Gaussian (f_dc) at feature_lr
            {"params": [self._features], "lr": feature_lr / 20.0, "name": "f_rest",
             "lead_lr": feature_lr, "lead": 3, "row": row},
            {"params": [self._opacity], "lr": opacity_lr, "name": "opacity"},
            {"params": [self._scaling], "lr": scaling_lr, "name": "scaling"},
            {"params": [self._rotation], "lr": rotation_lr, "name": "rotation"},
        ]
        self.optimizer = FusedAdam(groups, lr=0.0, eps=1e-15)
        return self.optimizer


    # -- adaptive density control (scene/gaussian_model_cf.py:452-457,523-719) --
    def add_densification_stats(self, viewspace_points, radii):
        from . import densify
        densify.add_densification_stats(self, viewspace_points, radii)

    def densify_and_prune(self, max_grad, min_opacity, extent, max_screen_size=None, generator=None):
        from . import densify
        return densify.densify_and_prune(self, max_grad, min_opacity, extent, max_screen_size, generator)

    def densify(self, max_grad, min_opacity, extent, max_screen_size=None, generator=None):
        from . import densify
        return densify.densify(self, max_grad, min_opacity, extent, max_screen_size, generator)

    def prune(self, max_grad, min_opacity, extent, max_screen_size=None):
        from . import densify
        return densify.prune(self, max_grad, min_opacity, extent, max_screen_size)

    def reset_opacity(self):
        from . import densify
        densify.reset_opacity(self)

    # -- PLY files (scene/gaussian_model_cf.py:398-432,459-515) --
    def save_ply(self, path: str) -> None:
        from . import ply
        ply.save_ply(path, self._xyz, self._features_dc, self._features_rest, self._opacity, self._scaling,
                     self._rotation)

    @classmethod
    def from_ply(cls, path: str, max_sh_degree: int = 3, num_gauss: int = -1, device="cuda") -> "GaussianParams":
        """load_ply: the parameters of a Gaussian PLY file, active SH degree
        = max_sh_degree."""
        from . import ply
        d = ply.load_ply(path, max_sh_degree, num_gauss)
        t = {k: torch.from_numpy(v) for k, v in d.items()}
        return cls(t["xyz"], t["features_dc"], t["features_rest"], t["scaling"], t["rotation"], t["opacity"],
                   max_sh_degree, device=device)


def render(gaussians: GaussianParams, cam: Camera, bg: torch.Tensor, pose=None, scaling_modifier: float = 1.0,
           clamp: bool = True, adam_in_backward: bool = False):
    """CF3DGS_Render.render with the rasteriser's 3-tuple return.  With
    ``clamp=False`` "image" is the rasteriser's colour before the [0, 1]
    clamp, for a loss that applies the clamp itself (the fused loss's
    ``clamp01``).  With ``adam_in_backward`` the backward applies the
    ``training_setup`` optimiser's step to every Gaussian parameter itself:
    the rasteriser's backward steps the means and SH coefficients, the
    activations' backward steps the raw scaling, rotation and opacity, and no
    gradient of theirs is written.  ``optimizer.step()`` then has nothing
    left to do for them."""
    means3D = gaussians.get_xyz
    screenspace = gaussians.screenspace_leaf()
    rs = raster_settings(cam, bg, gaussians.active_sh_degree, scaling_modifier)
    rasterizer = GaussianRasterizer(raster_settings=rs)
    fused = fused_act = None
    if adam_in_backward:
        if gaussians.optimizer is None:
            raise RuntimeError("adam_in_backward needs training_setup() first")
        opt = gaussians.optimizer
        fused = opt.in_backward(gaussians._xyz, gaussians._features)
        fused_act = opt.in_backward_activations(gaussians._scaling, gaussians._rotation, gaussians._opacity)
    scales, rotations, opacities = gaussians.get_activated(fused_act)
    color, radii, depth = rasterizer(means3D=means3D, means2D=screenspace, shs=gaussians.get_features,
                                     colors_precomp=None, opacities=opacities, scales=scales, rotations=rotations,
                                     cov3D_precomp=None, pose=pose, adam_in_backward=fused)
    return _RenderOut(image=color.clamp(0, 1) if clamp else color, depth=depth, viewspace_points=screenspace,
                      radii=radii)


class _RenderOut(dict):
    """The render dict of CF3DGS_Render.render; "visibility_filter"
    (radii > 0) is computed on first access, so a step that never reads it
    launches no compare kernel."""

    def __missing__(self, key):
        if key == "visibility_filter":
            v = self["radii"] > 0
            self[key] = v
            return v
        raise KeyError(key)
