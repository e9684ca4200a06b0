"""Camera matrices in CF-3DGS's conventions.

  * world-to-view ``getWorld2View2`` (utils/graphics_utils.py:57-68): R is the
    camera-to-world rotation, t the world-to-camera translation;
  * ``getProjectionMatrix`` (utils/graphics_utils.py:85-105);
  * the CF pipeline's camera, ``Camera(..., is_co3d=True)``
    (scene/cameras.py:57-66, built at trainer/trainer.py:709-719):
    ``getWorld2View3`` (utils/graphics_utils.py:71-82, R is the
    world-to-camera rotation) and a projection from the intrinsics with the
    principal point (cx, cy);
  * ``Camera`` (scene/cameras.py:51-74): world_view_transform = W2V^T,
    projection_matrix = P^T, full_proj_transform = W2V^T @ P^T,
    camera_center = inverse(W2V^T)[3, :3]; znear 0.01, zfar 100;
  * ``tanfov = tan(FoV / 2)`` as at scene/gaussian_model_cf.py:808-809.
Camera axes follow the reference: x right, y down, z forward.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


def fov2focal(fov: float, pixels: int) -> float:
    return pixels / (2 * math.tan(fov / 2))


def focal2fov(focal: float, pixels: int) -> float:
    return 2 * math.atan(pixels / (2 * focal))


def world_to_view(R: np.ndarray, t: np.ndarray, translate=(0.0, 0.0, 0.0), scale: float = 1.0) -> np.ndarray:
    Rt = np.zeros((4, 4))
    Rt[:3, :3] = np.asarray(R).transpose()
    Rt[:3, 3] = t
    Rt[3, 3] = 1.0
    C2W = np.linalg.inv(Rt)
    C2W[:3, 3] = (C2W[:3, 3] + np.asarray(translate)) * scale
    return np.float32(np.linalg.inv(C2W))


def projection_matrix(znear: float, zfar: float, fovx: float, fovy: float) -> torch.Tensor:
    tan_y = math.tan(fovy / 2)
    tan_x = math.tan(fovx / 2)
    top = tan_y * znear
    bottom = -top
    right = tan_x * znear
    left = -right
    P = torch.zeros(4, 4)
    P[0, 0] = 2.0 * znear / (right - left)
    P[1, 1] = 2.0 * znear / (top - bottom)
    P[0, 2] = (right + left) / (right - left)
    P[1, 2] = (top + bottom) / (top - bottom)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def world_to_view3(R: np.ndarray, t: np.ndarray, translate=(0.0, 0.0, 0.0), scale: float = 1.0) -> np.ndarray:
    """getWorld2View3: as world_to_view, but R enters untransposed (it is the
    world-to-camera rotation)."""
    Rt = np.zeros((4, 4))
    Rt[:3, :3] = np.asarray(R)
    Rt[:3, 3] = t
    Rt[3, 3] = 1.0
    C2W = np.linalg.inv(Rt)
    C2W[:3, 3] = (C2W[:3, 3] + np.asarray(translate)) * scale
    return np.float32(np.linalg.inv(C2W))


def intrinsics_projection(K, width: int, height: int, znear: float = 0.01, zfar: float = 100.0) -> torch.Tensor:
    """The is_co3d projection (scene/cameras.py:60-66), untransposed: from the
    float32 intrinsics fx = K[0,0], fy = K[1,1], cx = K[0,2], cy = K[1,2].
    A point (X, Y, Z) in the camera frame lands on the pixel centre
    (fx X / Z + cx - 0.5, fy Y / Z + cy - 0.5) of the rasteriser."""
    K = np.asarray(K, np.float32)
    fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
    w, h = width, height
    far, near = zfar, znear
    return torch.tensor([[2 * fx / w, 0.0, -(w - 2 * cx) / w, 0.0],
                         [0.0, 2 * fy / h, -(h - 2 * cy) / h, 0.0],
                         [0.0, 0.0, far / (far - near), -(far * near) / (far - near)],
                         [0.0, 0.0, 1.0, 0.0]]).float()


@dataclass
class Camera:
    image_width: int
    image_height: int
    FoVx: float
    FoVy: float
    world_view_transform: torch.Tensor
    projection_matrix: torch.Tensor
    full_proj_transform: torch.Tensor
    camera_center: torch.Tensor

    @property
    def tanfovx(self) -> float:
        return math.tan(self.FoVx * 0.5)

    @property
    def tanfovy(self) -> float:
        return math.tan(self.FoVy * 0.5)

    def to(self, device) -> "Camera":
        return Camera(self.image_width, self.image_height, self.FoVx, self.FoVy,
                      self.world_view_transform.to(device), self.projection_matrix.to(device),
                      self.full_proj_transform.to(device), self.camera_center.to(device))


def make_camera(R, t, fovx: float, fovy: float, width: int, height: int, znear: float = 0.01, zfar: float = 100.0,
                device="cpu") -> Camera:
    wv = torch.tensor(world_to_view(np.asarray(R, np.float64), np.asarray(t, np.float64))).transpose(0, 1)
    pm = projection_matrix(znear, zfar, fovx, fovy).transpose(0, 1)
    full = (wv.unsqueeze(0).bmm(pm.unsqueeze(0))).squeeze(0)
    center = wv.inverse()[3, :3]
    return Camera(width, height, fovx, fovy, wv.contiguous(), pm.contiguous(), full.contiguous(),
                  center.contiguous()).to(device)


def make_camera_co3d(R, t, K, width: int, height: int, znear: float = 0.01, zfar: float = 100.0,
                     device="cpu") -> Camera:
    """Camera(..., intrinsics=K, is_co3d=True) of the CF pipeline: R the
    world-to-camera rotation (identity and t = 0 for a new frame's pose solve,
    trainer/trainer.py:709-714), FoVx / FoVy = focal2fov of fx / fy as the
    data loaders set them (trainer/trainer.py:108-109)."""
    K = np.asarray(K, np.float32)
    wv = torch.tensor(world_to_view3(np.asarray(R, np.float64), np.asarray(t, np.float64))).transpose(0, 1)
    pm = intrinsics_projection(K, width, height, znear, zfar).transpose(0, 1)
    full = (wv.unsqueeze(0).bmm(pm.unsqueeze(0))).squeeze(0)
    center = wv.inverse()[3, :3]
    fovx = focal2fov(float(K[0, 0]), width)
    fovy = focal2fov(float(K[1, 1]), height)
    return Camera(width, height, fovx, fovy, wv.contiguous(), pm.contiguous(), full.contiguous(),
                  center.contiguous()).to(device)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """(R, t) for a camera at ``eye`` looking at ``target``; image up ~ ``up``."""
    eye = np.asarray(eye, np.float64)
    z = np.asarray(target, np.float64) - eye
    z /= np.linalg.norm(z)
    u = np.asarray(up, np.float64)
    y = -(u - np.dot(u, z) * z)
    if np.linalg.norm(y) < 1e-9:
        y = np.array([0.0, 0.0, 1.0]) - z[2] * z
    y /= np.linalg.norm(y)
    x = np.cross(y, z)
    R = np.stack([x, y, z], axis=1)  # camera-to-world rotation (columns = camera axes)
    t = -R.T @ eye
    return R, t


def raster_settings(cam: Camera, bg: torch.Tensor, sh_degree: int, scale_modifier: float = 1.0,
                    prefiltered: bool = False, debug: bool = False):
    """GaussianRasterizationSettings exactly as CF3DGS_Render.render builds it."""
    from diff_gaussian_rasterization import GaussianRasterizationSettings
    return GaussianRasterizationSettings(
        image_height=int(cam.image_height), image_width=int(cam.image_width),
        tanfovx=cam.tanfovx, tanfovy=cam.tanfovy, bg=bg, scale_modifier=scale_modifier,
        viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform, sh_degree=sh_degree,
        campos=cam.camera_center, prefiltered=prefiltered, debug=debug)


def oracle_camera(cam: Camera, bg, sh_degree: int, scale_modifier: float = 1.0, pose=None) -> dict:
    """The same camera as plain numpy arrays (for the CPU oracle)."""
    d = dict(W=cam.image_width, H=cam.image_height, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
             view=cam.world_view_transform.detach().cpu().numpy().reshape(-1),
             proj=cam.full_proj_transform.detach().cpu().numpy().reshape(-1),
             campos=cam.camera_center.detach().cpu().numpy().reshape(-1),
             bg=np.asarray(bg.detach().cpu().numpy() if isinstance(bg, torch.Tensor) else bg, np.float32).reshape(-1),
             sh_degree=sh_degree, scale_modifier=scale_modifier)
    if pose is not None:
        d["pose7"] = np.asarray(pose.detach().cpu().numpy() if isinstance(pose, torch.Tensor) else pose,
                                np.float32).reshape(-1)
    return d
