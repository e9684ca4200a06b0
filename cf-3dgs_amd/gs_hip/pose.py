"""Per-frame pose solve of CF-3DGS on the gfx950 rasteriser.

The reference optimises, for every new frame, only the rigid pose of the local
Gaussians (trainer/cf3dgs_trainer.py:261-285): 300 iterations of
render -> image.clamp(0, 1) -> (0.8 L1 + 0.2 (1 - SSIM)) -> backward ->
Adam(lr = rotation_lr = 1e-3, eps = 1e-15) on the lietorch
``LieGroupParameter`` P[0] (scene/gaussian_model_cf.py:303-340), whose data is
a 6-D tangent xi (zero at start) with the pose Exp(xi) * X0, X0 = identity
(init_RT(None)).  The pose moves the means only (``P[0].retr().act(xyz)``,
:123-133); here the renderer applies it inside preprocess.

This module holds the instance builder (the frame-ingest of
trainer/trainer.py:476-512 + create_from_pcd, scene/gaussian_model_cf.py:178-212,
on a synthetic depth map), the error metrics of
utils/utils_poses/comp_ate.py:13-38, and the *plain* solver: one Python
iteration per step through the drop-in ``GaussianRasterizer`` with torch ops
around it, as the reference drives it.  The graph-captured batched solver is
in ``batch.py``.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from . import raster as _raster
from .camera import focal2fov, fov2focal, make_camera, raster_settings
from .loss import photometric_loss
from .scenes import inverse_sigmoid, random_pose, rgb2sh, synthetic_depth

LR = 1e-3      # rotation_lr, arguments/__init__.py:111
EPS = 1e-15    # torch.optim.Adam(..., eps=1e-15), scene/gaussian_model_cf.py:322
ITERS = 300    # optim_opt.iterations for a new frame, trainer/cf3dgs_trainer.py:264


# ---------------------------------------------------------------------------
# SE(3) in torch (lietorch conventions: tangent [tau, phi], pose [t, q_xyzw])

def _cross(a, b):
    return torch.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                        a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], dim=-1)


def quat_mul(q1, q2):
    """Hamilton product of [x, y, z, w] quaternions."""
    u1, w1 = q1[..., :3], q1[..., 3:]
    u2, w2 = q2[..., :3], q2[..., 3:]
    u = w1 * u2 + w2 * u1 + _cross(u1, u2)
    w = w1 * w2 - (u1 * u2).sum(-1, keepdim=True)
    return torch.cat([u, w], dim=-1)


def quat_rotate(q, v):
    u, w = q[..., :3], q[..., 3:]
    uv = _cross(u, v)
    return v + 2.0 * w * uv + 2.0 * _cross(u, uv)


def se3_exp(xi: torch.Tensor, X0: Optional[torch.Tensor] = None) -> torch.Tensor:
    """[..., 7] pose = Exp(xi) * X0 (X0 = identity when None), differentiable
    everywhere including xi = 0 (Taylor branch in th^2 for small angles)."""
    tau, phi = xi[..., :3], xi[..., 3:]
    th2 = (phi * phi).sum(-1, keepdim=True)
    small = th2 < 1e-6
    th2s = torch.where(small, torch.ones_like(th2), th2)  # keeps the unused branch finite
    th = torch.sqrt(th2s)
    s_big = torch.sin(0.5 * th) / th
    w_big = torch.cos(0.5 * th)
    c1_big = (1.0 - torch.cos(th)) / th2s
    c2_big = (th - torch.sin(th)) / (th2s * th)
    th4 = th2 * th2
    s = torch.where(small, 0.5 - th2 / 48.0 + th4 / 3840.0, s_big)
    w = torch.where(small, 1.0 - th2 / 8.0 + th4 / 384.0, w_big)
    c1 = torch.where(small, 0.5 - th2 / 24.0 + th4 / 720.0, c1_big)
    c2 = torch.where(small, 1.0 / 6.0 - th2 / 120.0 + th4 / 5040.0, c2_big)
    qa = torch.cat([s * phi, w], dim=-1)
    k1 = _cross(phi, tau)
    ta = tau + c1 * k1 + c2 * _cross(phi, k1)
    if X0 is None:
        return torch.cat([ta, qa], dim=-1)
    t0, q0 = X0[..., :3], X0[..., 3:]
    return torch.cat([ta + quat_rotate(qa, t0), quat_mul(qa, q0)], dim=-1)


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    x, y, z, w = (float(v) for v in q / np.linalg.norm(q))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def pose_matrix(pose7) -> np.ndarray:
    p = np.asarray(torch.as_tensor(pose7).detach().cpu().double())
    T = np.eye(4)
    T[:3, :3] = quat_to_matrix(p[3:7])
    T[:3, 3] = p[:3]
    return T


def pose_errors(pose_est, pose_ref):
    """(rotation error [rad], translation error) of E = T_ref^-1 T_est, as
    rotation_error / translation_error of utils/utils_poses/comp_ate.py:13-38."""
    E = np.linalg.inv(pose_matrix(pose_ref)) @ pose_matrix(pose_est)
    d = 0.5 * (E[0, 0] + E[1, 1] + E[2, 2] - 1.0)
    rot = float(np.arccos(max(min(d, 1.0), -1.0)))
    trans = float(np.sqrt((E[:3, 3] ** 2).sum()))
    return rot, trans


# ---------------------------------------------------------------------------
# one frame of the pose-solve workload

@dataclass
class PoseInstance:
    """The local Gaussians of one frame (create_from_pcd on its lifted depth),
    its camera, the target image (a render at the hidden pose ``gt_pose``)."""
    seed: int
    W: int
    H: int
    P: int
    means: torch.Tensor       # [P, 3]
    shs: torch.Tensor         # [P, 1, 3]  (active SH degree 0 reads only the DC term)
    opacities: torch.Tensor   # [P, 1]  sigmoid(inverse_sigmoid(0.1))
    scales: torch.Tensor      # [P, 3]  exp(log(sqrt(clamp(distCUDA2, 1e-7))))
    rotations: torch.Tensor   # [P, 4]  (1, 0, 0, 0)
    settings: object          # GaussianRasterizationSettings (view = identity)
    target: torch.Tensor      # [3, H, W] in [0, 1]
    gt_pose: torch.Tensor     # [7] device
    supergrid: bool = False


def lift_points(seed: int, W: int, H: int, P: Optional[int], fovx_deg: float = 60.0):
    """Depth lift (X = (u - cx) d / fx, Y = (v - cy) d / fy, Z = d, the
    kornia depth_to_3d semantics of trainer/trainer.py:489-495) of a seeded
    synthetic depth map, subsampled to P points without replacement; when P
    exceeds W*H the map is lifted on a 2x super-grid (2W x 2H samples at
    half-pixel steps, same camera)."""
    g = torch.Generator().manual_seed(seed)
    fx = fov2focal(math.radians(fovx_deg), W)
    fy = fx
    cx, cy = W / 2.0, H / 2.0
    supergrid = P is not None and P > W * H
    k = 2 if supergrid else 1
    d = synthetic_depth(W * k, H * k, g)
    v, u = torch.meshgrid(torch.arange(H * k, dtype=torch.float32) / k,
                          torch.arange(W * k, dtype=torch.float32) / k, indexing="ij")
    xyz = torch.stack([(u - cx) * d / fx, (v - cy) * d / fy, d], dim=-1).reshape(-1, 3)
    rgb = torch.rand((xyz.shape[0], 3), generator=g)
    if P is not None and P != xyz.shape[0]:
        idx = torch.randperm(xyz.shape[0], generator=g)[:P]
        xyz, rgb = xyz[idx], rgb[idx]
    fovy = focal2fov(fy, H)
    cam = make_camera(np.eye(3), np.zeros(3), math.radians(fovx_deg), fovy, W, H)
    return xyz.contiguous(), rgb.contiguous(), cam, supergrid


def build_instance(seed: int, W: int = 480, H: int = 270, P: Optional[int] = None, fovx_deg: float = 60.0,
                   device="cuda", sigma_rot: float = 0.05, sigma_t: float = 0.05) -> PoseInstance:
    """One frame: lift, create_from_pcd (scene/gaussian_model_cf.py:178-212:
    RGB2SH colours, scale log(sqrt(clamp(distCUDA2, 1e-7))) x 3, rotation
    (1,0,0,0), opacity inverse_sigmoid(0.1)), activations applied
    (get_scaling exp, get_opacity sigmoid), hidden pose from ``random_pose``
    (axis-angle and translation ~ N(0, 0.05^2 I)), target = render at it."""
    device = torch.device(device)
    xyz, rgb, cam, supergrid = lift_points(seed, W, H, P, fovx_deg)
    means = xyz.to(device)
    dist2 = torch.clamp_min(_raster.knn3_meandist(means), 1e-7)
    log_scale = torch.log(torch.sqrt(dist2))[:, None].repeat(1, 3)
    n = means.shape[0]
    rots = torch.zeros((n, 4), device=device)
    rots[:, 0] = 1.0
    opac_raw = inverse_sigmoid(0.1 * torch.ones((n, 1), device=device))
    shs = rgb2sh(rgb).to(device).reshape(n, 1, 3).contiguous()
    bg = torch.zeros(3, device=device)
    rs = raster_settings(cam.to(device), bg, 0)
    inst = PoseInstance(seed=seed, W=W, H=H, P=n, means=means, shs=shs, opacities=torch.sigmoid(opac_raw),
                        scales=torch.exp(log_scale), rotations=rots / rots.norm(dim=-1, keepdim=True),
                        settings=rs, target=None, gt_pose=random_pose(seed + 7919, sigma_rot, sigma_t).to(device),
                        supergrid=supergrid)
    with torch.no_grad():
        color, _, _, _, _ = _raster.forward(rs, inst.means, inst.shs, None, inst.opacities, inst.scales,
                                            inst.rotations, None, pose=inst.gt_pose, want_alpha=False)
    inst.target = color.clamp(0.0, 1.0).contiguous()
    return inst


def frame_payload(seed: int, W: int = 480, H: int = 270, fovx_deg: float = 60.0):
    """What a pose-solve request brings: the new frame's depth map [H, W],
    its RGB image [H, W, 3] in [0, 1] (the seeded draws of ``lift_points``,
    pixel for pixel) and its intrinsics (fx, fy, cx, cy), plus the frame's
    camera (identity view)."""
    g = torch.Generator().manual_seed(seed)
    fx = fov2focal(math.radians(fovx_deg), W)
    d = synthetic_depth(W, H, g)
    rgb = torch.rand((H * W, 3), generator=g).reshape(H, W, 3)
    cam = make_camera(np.eye(3), np.zeros(3), math.radians(fovx_deg), focal2fov(fx, H), W, H)
    return d, rgb, (fx, fx, W / 2.0, H / 2.0), cam


def build_instance_device(seed: int, W: int = 480, H: int = 270, fovx_deg: float = 60.0, device="cuda",
                          sigma_rot: float = 0.05, sigma_t: float = 0.05) -> PoseInstance:
    """``build_instance`` with the frame ingested on the device: the request's
    depth map and image are uploaded and go through the library's frame
    ingest (lift, distCUDA2 scales, create_from_pcd with one Gaussian per
    pixel and active SH degree 0, ``gs_hip.ingest``), then the activations
    and the target render at the hidden pose.  Only the payload crosses from
    the host."""
    from . import ingest
    device = torch.device(device)
    d, rgb, K, cam = frame_payload(seed, W, H, fovx_deg)
    r = ingest.ingest_frame(d.to(device), rgb.to(device), K, voxel_size=None, max_sh_degree=0,
                            channels_first=False)
    n = r["xyz"].shape[0]
    rs = raster_settings(cam.to(device), torch.zeros(3, device=device), 0)
    inst = PoseInstance(seed=seed, W=W, H=H, P=n, means=r["xyz"], shs=r["features"].contiguous(),
                        opacities=torch.sigmoid(r["opacity"]), scales=torch.exp(r["scaling"]),
                        rotations=torch.nn.functional.normalize(r["rotation"], dim=-1), settings=rs, target=None,
                        gt_pose=random_pose(seed + 7919, sigma_rot, sigma_t).to(device))
    with torch.no_grad():
        color, _, _, _, _ = _raster.forward(rs, inst.means, inst.shs, None, inst.opacities, inst.scales,
                                            inst.rotations, None, pose=inst.gt_pose, want_alpha=False)
    inst.target = color.clamp(0.0, 1.0).contiguous()
    return inst


# ---------------------------------------------------------------------------
# plain solver: the reference's loop shape, one iteration per Python step

def solve_plain(inst: PoseInstance, iters: int = ITERS, lr: float = LR, lambda_dssim: float = 0.2,
                X0: Optional[torch.Tensor] = None, fused_loss: bool = False):
    """Returns dict(pose [7], xi [6], loss, seconds).

    Per iteration, like train_step (trainer/cf3dgs_trainer.py:81-173) with
    only the pose optimiser stepping: pose = Exp(xi) * X0 in torch ops, the
    drop-in rasteriser (forward readback of num_rendered included),
    image.clamp(0, 1), the torch loss (or the fused HIP loss), backward,
    the ``torch.isnan(loss)`` check (trainer/losses.py:126, a host sync), and
    torch.optim.Adam on xi."""
    from diff_gaussian_rasterization import GaussianRasterizer
    device = inst.means.device
    xi = torch.zeros(6, device=device, requires_grad=True)
    x0 = X0.to(device) if X0 is not None else None
    opt = torch.optim.Adam([xi], lr=lr, eps=EPS)
    rasterizer = GaussianRasterizer(raster_settings=inst.settings)
    means2D = torch.zeros_like(inst.means)
    loss = None
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(iters):
        pose = se3_exp(xi, x0)
        color, radii, depth = rasterizer(means3D=inst.means, means2D=means2D, opacities=inst.opacities,
                                         shs=inst.shs, scales=inst.scales, rotations=inst.rotations, pose=pose)
        if fused_loss:
            from .loss import fused_photometric_loss
            loss, _ = fused_photometric_loss(color, inst.target, lambda_dssim, clamp01=True)
        else:
            loss = photometric_loss(color.clamp(0.0, 1.0), inst.target, lambda_dssim)
        loss.backward()
        if torch.isnan(loss):
            raise FloatingPointError("pose solve: loss is NaN")
        opt.step()
        opt.zero_grad(set_to_none=True)
    torch.cuda.synchronize(device)
    seconds = time.perf_counter() - t0
    with torch.no_grad():
        pose = se3_exp(xi, x0)
    return dict(pose=pose.detach(), xi=xi.detach().clone(),
                loss=float(loss.detach()) if loss is not None else float("nan"),
                seconds=seconds)


def solve_plain_native(inst: PoseInstance, iters: int = ITERS, lr: float = LR, lambda_dssim: float = 0.2,
                       X0: Optional[torch.Tensor] = None, pose_only: bool = False):
    """The plain one-solve-per-launch path on this library's kernels: per
    iteration an eager launch sequence (preprocess, num_rendered readback,
    binning + blend, fused loss forward and backward, the rasteriser backward
    with fixed-order reductions, SE(3) Adam) and the reference's NaN check on
    the host.  The backward computes every gradient, as the reference's
    rasteriser does for a pose solve (its Gaussians' parameters still require
    gradients); ``pose_only`` asks for the pose gradient alone, which runs
    the pose-only backward.  Every reduction is in a fixed order, so either
    way it computes bit for bit what the batched solver computes for the same
    instance.  Returns dict(pose [7], loss, seconds)."""
    from .loss import fused_photometric_value_and_grad
    device = inst.means.device
    opt = SE3Adam(1, device, lr=lr, eps=EPS, X0=X0)
    stats = None
    want = {"pose"} if pose_only else None
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(iters):
        pose = opt.pose[0]
        color, _, _, _, st = _raster.forward(inst.settings, inst.means, inst.shs, None, inst.opacities, inst.scales,
                                             inst.rotations, None, pose=pose, want_alpha=False)
        stats, dcolor = fused_photometric_value_and_grad(color, inst.target, lambda_dssim, clamp01=True)
        g = _raster.backward(inst.settings, st, inst.means, inst.shs, None, inst.opacities, inst.scales,
                             inst.rotations, None, pose, dcolor, want=want, deterministic=True)
        if torch.isnan(stats[0]):
            raise FloatingPointError("pose solve: loss is NaN")
        opt.grad[0].copy_(g["pose"])
        opt.step()
    torch.cuda.synchronize(device)
    seconds = time.perf_counter() - t0
    return dict(pose=opt.pose[0].clone(), loss=float(stats[0]) if stats is not None else float("nan"),
                seconds=seconds)


# ---------------------------------------------------------------------------
# C-ABI wrappers of the SE(3) kernels

def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def se3_retract(xi: torch.Tensor, X0: Optional[torch.Tensor] = None) -> torch.Tensor:
    """[n, 7] = Exp(xi) * X0 on the device (gs_se3_retract)."""
    lib = _lib.load()
    x = xi.reshape(-1, 6).contiguous().float()
    n = x.shape[0]
    out = torch.empty((n, 7), dtype=torch.float32, device=x.device)
    x0 = X0.reshape(-1, 7).contiguous().float() if X0 is not None else None
    _lib.check(lib.gs_se3_retract(n, _p(x), _p(x0), _p(out), _raster._stream(x.device)), "gs_se3_retract")
    return out


def se3_grad(xi: torch.Tensor, grad_pose: torch.Tensor, X0: Optional[torch.Tensor] = None) -> torch.Tensor:
    """[n, 6] = (d pose / d xi)^T grad_pose on the device (gs_se3_grad)."""
    lib = _lib.load()
    x = xi.reshape(-1, 6).contiguous().float()
    g = grad_pose.reshape(-1, 7).contiguous().float()
    n = x.shape[0]
    out = torch.empty((n, 6), dtype=torch.float32, device=x.device)
    x0 = X0.reshape(-1, 7).contiguous().float() if X0 is not None else None
    _lib.check(lib.gs_se3_grad(n, _p(x), _p(x0), _p(g), _p(out), _raster._stream(x.device)), "gs_se3_grad")
    return out


class SE3Adam:
    """Adam over n tangent parameters in one launch (gs_se3_adam): the device
    step count makes the update replayable inside a captured graph."""

    def __init__(self, n: int, device, lr: float = LR, betas=(0.9, 0.999), eps: float = EPS,
                 X0: Optional[torch.Tensor] = None):
        self.n = n
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.xi = torch.zeros((n, 6), dtype=torch.float32, device=device)
        self.m = torch.zeros_like(self.xi)
        self.v = torch.zeros_like(self.xi)
        self.step_count = torch.zeros(1, dtype=torch.int32, device=device)
        self.X0 = X0.reshape(n, 7).contiguous().float().to(device) if X0 is not None else None
        self.pose = torch.empty((n, 7), dtype=torch.float32, device=device)
        self.grad = torch.zeros((n, 7), dtype=torch.float32, device=device)
        self.reset()

    def reset(self):
        """xi = 0, moments = 0, step = 0, pose = X0 (filled by kernels, so a
        reset may also be captured)."""
        self.xi.zero_()
        self.m.zero_()
        self.v.zero_()
        self.step_count.zero_()
        lib = _lib.load()
        _lib.check(lib.gs_se3_retract(self.n, _p(self.xi), _p(self.X0), _p(self.pose),
                                      _raster._stream(self.xi.device)), "gs_se3_retract")

    def step(self):
        lib = _lib.load()
        _lib.check(lib.gs_se3_adam(self.n, _p(self.xi), _p(self.X0), _p(self.grad), _p(self.m), _p(self.v),
                                   _p(self.step_count), self.lr, self.b1, self.b2, self.eps, _p(self.pose),
                                   _raster._stream(self.xi.device)), "gs_se3_adam")
