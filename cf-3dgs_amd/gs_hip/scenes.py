"""Seeded synthetic workloads (CPU torch generators; move to the GPU after).

``plain_scene``      10k Gaussians, 256x256, 60 deg FoV, means U[-1,1]^3 at
                     distance 4, log-scales N(-3, 0.5^2), normalised N(0,I4)
                     quaternions, sigmoid(N(0,1)) opacities, SH degree 3 with
                     N(0, 0.1^2) coefficients, black background; L1 target
                     U[0,1] from its own seed.
``training_scene``   500k Gaussians in 32 isotropic clusters (sigma 0.3) with
                     centres U([-3,3]^3); 960x540 camera 10 units from the
                     centroid (on -z), FoVx 60 deg; raw (pre-activation)
                     parameters as CFGaussianModel stores them
                     (scene/gaussian_model_cf.py:41-56, 178-212).
``depth_lift_frame`` one Gaussian per pixel lifted from a smooth synthetic
                     depth map (2 + 0.5 * normalised sum of 8 sinusoids), RGB
                     noise colours as SH DC, isotropic kNN scale, opacity 0.1,
                     identity rotation -- create_from_pcd
                     (scene/gaussian_model_cf.py:178-212) on a lifted frame.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .camera import focal2fov, fov2focal, look_at, make_camera

SH_C0 = 0.28209479177387814


def rgb2sh(rgb):
    """utils/sh_utils.py:114-115."""
    return (rgb - 0.5) / SH_C0


def inverse_sigmoid(x):
    """utils/general_utils.py:16-17."""
    return torch.log(x / (1 - x))


def plain_scene(seed: int = 0, P: int = 10000, W: int = 256, H: int = 256, target_seed: int = 1):
    g = torch.Generator().manual_seed(seed)
    means = torch.rand((P, 3), generator=g) * 2 - 1
    means[:, 2] += 4.0
    log_s = torch.randn((P, 3), generator=g) * 0.5 - 3.0
    q = torch.randn((P, 4), generator=g)
    q = q / q.norm(dim=1, keepdim=True)
    opac = torch.sigmoid(torch.randn((P, 1), generator=g))
    shs = torch.randn((P, 16, 3), generator=g) * 0.1
    fov = math.radians(60.0)
    cam = make_camera(np.eye(3), np.zeros(3), fov, fov, W, H)
    gt = torch.Generator().manual_seed(target_seed)
    target = torch.rand((3, H, W), generator=gt)
    return dict(means=means, scales=torch.exp(log_s), rotations=q, opacities=opac, shs=shs, cam=cam,
                bg=torch.zeros(3), sh_degree=3, target=target)


def training_scene(seed: int = 0, P: int = 500_000, W: int = 960, H: int = 540, n_clusters: int = 32,
                   sigma: float = 0.3, distance: float = 10.0, target_seed: int = 1):
    g = torch.Generator().manual_seed(seed)
    centres = torch.rand((n_clusters, 3), generator=g) * 6 - 3
    assign = torch.randint(0, n_clusters, (P,), generator=g)
    xyz = centres[assign] + torch.randn((P, 3), generator=g) * sigma
    log_s = torch.randn((P, 3), generator=g) * 0.5 - 3.0
    rot = torch.randn((P, 4), generator=g)
    opac_logit = torch.randn((P, 1), generator=g)
    sh = torch.randn((P, 16, 3), generator=g) * 0.1
    centroid = xyz.mean(dim=0).double().numpy()
    fovx = math.radians(60.0)
    fovy = focal2fov(fov2focal(fovx, W), H)
    R, t = look_at(centroid + np.array([0.0, 0.0, -distance]), centroid)
    cam = make_camera(R, t, fovx, fovy, W, H)
    gt = torch.Generator().manual_seed(target_seed)
    target = torch.rand((3, H, W), generator=gt)
    return dict(xyz=xyz, features_dc=sh[:, :1, :].contiguous(), features_rest=sh[:, 1:, :].contiguous(),
                scaling=log_s, rotation=rot, opacity=opac_logit, cam=cam, bg=torch.zeros(3), sh_degree=3,
                target=target)


def synthetic_depth(W: int, H: int, g: torch.Generator) -> torch.Tensor:
    """2 + 0.5 * s / max|s|, s = sum of 8 sinusoids with periods 32..256 px."""
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64),
                          indexing="ij")
    s = torch.zeros((H, W), dtype=torch.float64)
    for _ in range(8):
        period = 32 + torch.rand((), generator=g, dtype=torch.float64) * (256 - 32)
        ang = torch.rand((), generator=g, dtype=torch.float64) * 2 * math.pi
        phase = torch.rand((), generator=g, dtype=torch.float64) * 2 * math.pi
        om = 2 * math.pi / period
        s += torch.sin(om * (math.cos(ang) * u + math.sin(ang) * v) + phase)
    return (2.0 + 0.5 * s / s.abs().max()).float()


def depth_lift_frame(seed: int, W: int = 480, H: int = 270, fovx_deg: float = 60.0, n_points=None):
    """Lifted point cloud of one synthetic frame (kornia depth_to_3d semantics,
    trainer/trainer.py:489-495): X = (u - cx) d / fx, Y = (v - cy) d / fy, Z = d.
    Returns xyz [N,3], rgb [N,3], the camera, fx."""
    g = torch.Generator().manual_seed(seed)
    d = synthetic_depth(W, H, g)
    fx = fov2focal(math.radians(fovx_deg), W)
    fy = fx
    cx, cy = W / 2.0, H / 2.0
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float32), torch.arange(W, dtype=torch.float32), indexing="ij")
    X = (u - cx) * d / fx
    Y = (v - cy) * d / fy
    xyz = torch.stack([X, Y, d], dim=-1).reshape(-1, 3)
    rgb = torch.rand((H * W, 3), generator=g)
    if n_points is not None and n_points != xyz.shape[0]:
        idx = torch.randperm(xyz.shape[0], generator=g)[:n_points]
        xyz, rgb = xyz[idx], rgb[idx]
    fovy = focal2fov(fy, H)
    cam = make_camera(np.eye(3), np.zeros(3), math.radians(fovx_deg), fovy, W, H)
    return dict(xyz=xyz.contiguous(), rgb=rgb.contiguous(), cam=cam, fx=fx)


def random_pose(seed: int, sigma_rot: float = 0.05, sigma_t: float = 0.05) -> torch.Tensor:
    """[tx, ty, tz, qx, qy, qz, qw]: axis-angle ~ N(0, sigma_rot^2 I), t ~ N(0, sigma_t^2 I)."""
    g = torch.Generator().manual_seed(seed)
    aa = torch.randn(3, generator=g, dtype=torch.float64) * sigma_rot
    t = torch.randn(3, generator=g, dtype=torch.float64) * sigma_t
    ang = float(aa.norm())
    if ang < 1e-12:
        q = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=torch.float64)
    else:
        axis = aa / ang
        q = torch.cat([axis * math.sin(ang / 2), torch.tensor([math.cos(ang / 2)], dtype=torch.float64)])
    return torch.cat([t, q]).float()
