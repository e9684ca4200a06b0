"""Frame ingest on the device: depth lift, voxel downsample, create_from_pcd.

The reference turns every new frame into local Gaussians on the host
(trainer/trainer.py:334-423 ``prepare_data_co3d``, :425-514
``prepare_data_from_viewpoint``).  It runs these steps:

  * kornia ``depth_to_3d`` lifts the depth on the device;
  * the cloud is copied to numpy and put into an open3d ``PointCloud``, which
    is reduced with ``voxel_down_sample(voxel_size=0.01)``;
  * ``create_from_pcd`` (scene/gaussian_model_cf.py:178-212) uploads the
    cloud again and calls ``distCUDA2`` for the initial scales.

Here every step is a kernel of libgs_hip.so.  The host reads back two small
values: the cloud's bounds (6 floats, which size the voxel key; open3d reads
GetMinBound / GetMaxBound at the same point) and the voxel count (which sizes
the parameter tensors).

Semantics kept from the reference:

  * the lift computes ``((u - cx) / fx) * d``, ``((v - cy) / fy) * d``, ``d``
    in float32, in kornia's ``unproject_points`` order;
  * 8-bit colours become ``v / 255.0`` in double, as ``np.asarray(image) /
    255.0``;
  * voxel indices are ``floor((p - (min - s/2)) / s)`` in double;
  * voxel means are double sums in point order, divided by the count, then
    rounded to float32 (``np.asarray(pcd.points, dtype=np.float32)``).

Voxels come out in ascending (ix, iy, iz) order.  open3d emits them in the
order of its hash map, which is unspecified.  ``estimate_normals`` is not
reproduced: ``create_from_pcd`` never reads the normals.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib
from .raster import _ptr, _stream, knn3_meandist, require_device

INT_MAX = 2 ** 31 - 1
VOXEL_SIZE = 0.01  # trainer/trainer.py:415,507


def _intrinsics(intrinsics) -> Tuple[float, float, float, float]:
    """(fx, fy, cx, cy) from a 3x3 matrix (tensor / array) or a 4-sequence.
    The values are taken as float32, as ``intr_mat_tensor = ...float()``."""
    if isinstance(intrinsics, torch.Tensor):
        intrinsics = intrinsics.detach().cpu().double().numpy()
    a = np.asarray(intrinsics, dtype=np.float64)
    if a.shape == (3, 3):
        fx, fy, cx, cy = a[0, 0], a[1, 1], a[0, 2], a[1, 2]
    elif a.shape == (4,):
        fx, fy, cx, cy = a
    else:
        raise ValueError("intrinsics must be a 3x3 matrix or (fx, fy, cx, cy)")
    return tuple(float(np.float32(x)) for x in (fx, fy, cx, cy))


def _color_source(colors: torch.Tensor, n: int, channels_first: Optional[bool] = None):
    """(tensor, is_u8, pixel stride, channel stride) for per-point colours:
    uint8 or float32, [N, 3], [H, W, 3] or [3, H, W]."""
    if colors.dtype == torch.uint8:
        u8 = 1
    else:
        u8 = 0
        if colors.dtype != torch.float32:
            colors = colors.float()
    colors = colors.contiguous()
    if colors.numel() != 3 * n:
        raise ValueError(f"colours hold {colors.numel()} values, expected 3 per point ({3 * n})")
    if channels_first is None:
        channels_first = colors.dim() == 3 and colors.shape[0] == 3 and colors.shape[-1] != 3
    if channels_first:
        return colors, u8, 1, n
    return colors, u8, 3, 1


def depth_to_3d(depth: torch.Tensor, intrinsics, image: Optional[torch.Tensor] = None,
                channels_first: Optional[bool] = None):
    """kornia ``depth_to_3d(depth[None, None], K[None], normalize_points=False)``
    followed by ``permute(1, 2, 0).reshape(-1, 3)`` (trainer/trainer.py:489-495).
    Returns the [H*W, 3] float32 points in row-major pixel order.  With
    ``image``, it also returns the per-point colours as float32 (8-bit values
    become v / 255 in double, then float32)."""
    device = require_device(depth, "depth")
    if depth.dim() == 3 and depth.shape[0] == 1:
        depth = depth[0]
    if depth.dim() != 2:
        raise ValueError("depth must be [H, W] or [1, H, W]")
    H, W = depth.shape
    d = depth.contiguous().float()
    fx, fy, cx, cy = _intrinsics(intrinsics)
    pts = torch.empty((H * W, 3), dtype=torch.float32, device=device)
    lib = _lib.load()
    cols = None
    src, u8, ps, cs = None, 0, 0, 0
    if image is not None:
        if image.device != device:
            raise ValueError(f"image is on {image.device}, expected {device}")
        src, u8, ps, cs = _color_source(image, H * W, channels_first)
        cols = torch.empty((H * W, 3), dtype=torch.float32, device=device)
    _lib.check(lib.gs_depth_to_points(W, H, _ptr(d), fx, fy, cx, cy, _ptr(pts), _ptr(src), u8, ps, cs, _ptr(cols),
                                      _stream(device)), "gs_depth_to_points")
    return pts if image is None else (pts, cols)


def points_bounds(points: torch.Tensor) -> Tuple[np.ndarray, np.ndarray]:
    """(min xyz, max xyz) of a [N, 3] float32 device cloud, read back to the host.
    Raises on a non-finite coordinate."""
    device = require_device(points, "points")
    n = points.shape[0]
    if n == 0:
        raise ValueError("empty point cloud")
    lib = _lib.load()
    ws = torch.empty(int(lib.gs_bounds_workspace_bytes()), dtype=torch.uint8, device=device)
    b = torch.empty(6, dtype=torch.float32, device=device)
    _lib.check(lib.gs_points_bounds(_ptr(points), n, _ptr(b), _ptr(ws), ws.numel(), _stream(device)),
               "gs_points_bounds")
    h = b.cpu().numpy().astype(np.float64)
    if not np.all(np.isfinite(h)):
        raise ValueError("point cloud holds a non-finite coordinate")
    return h[:3], h[3:]


def _bits_for(n: int) -> int:
    """Bits that hold the values 0 .. n-1."""
    return max(0, int(n - 1).bit_length())


def voxel_grid(mins: Sequence[float], maxs: Sequence[float], voxel_size: float):
    """open3d VoxelDownSample's grid (double precision): voxel_min_bound =
    min - s/2, the "voxel_size is too small" check against INT_MAX, and the
    bits per axis that hold the largest index floor((max - vmin) / s).
    Returns (vmin [3] float64, (bits_x, bits_y, bits_z))."""
    s = float(voxel_size)
    if not s > 0.0:
        raise ValueError("voxel_size <= 0.")
    mn = np.asarray(mins, dtype=np.float64)
    mx = np.asarray(maxs, dtype=np.float64)
    half = s * 0.5
    vmin = mn - half
    vmax = mx + half
    if s * INT_MAX < float(np.max(vmax - vmin)):
        raise ValueError("voxel_size is too small.")
    top = [int(math.floor((float(mx[a]) - float(vmin[a])) / s)) for a in range(3)]
    bits = tuple(_bits_for(t + 1) for t in top)
    if sum(bits) > 64:
        raise ValueError(f"voxel grid needs {sum(bits)} key bits (at most 64)")
    return vmin, bits


def voxel_down_sample(points: torch.Tensor, colors: torch.Tensor, voxel_size: float = VOXEL_SIZE,
                      channels_first: Optional[bool] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """open3d ``PointCloud.voxel_down_sample`` on the device: the mean point and
    colour of every occupied voxel (trainer/trainer.py:414-415, 506-507).
    ``colors`` are per point (uint8 v -> v / 255, or float32), as [N, 3],
    [H, W, 3] or [3, H, W] when the points are a lifted frame."""
    device = require_device(points, "points")
    pts = points.contiguous().float()
    if pts.dim() != 2 or pts.shape[1] != 3:
        raise ValueError("points must be [N, 3]")
    n = pts.shape[0]
    if colors.device != device:
        raise ValueError(f"colors are on {colors.device}, expected {device}")
    src, u8, ps, cs = _color_source(colors, n, channels_first)
    mins, maxs = points_bounds(pts)
    vmin, (bx, by, bz) = voxel_grid(mins, maxs, voxel_size)
    lib = _lib.load()
    ws = torch.empty(int(lib.gs_voxel_workspace_bytes(n)), dtype=torch.uint8, device=device)
    out_p = torch.empty((n, 3), dtype=torch.float32, device=device)
    out_c = torch.empty((n, 3), dtype=torch.float32, device=device)
    n_out = torch.zeros(1, dtype=torch.int32, device=device)
    vm = (ctypes.c_double * 3)(*[float(x) for x in vmin])
    _lib.check(lib.gs_voxel_down_sample(_ptr(pts), n, _ptr(src), u8, ps, cs, vm, float(voxel_size), bx, by, bz,
                                        _ptr(out_p), _ptr(out_c), _ptr(n_out), _ptr(ws), ws.numel(),
                                        _stream(device)), "gs_voxel_down_sample")
    m = int(n_out.item())
    return out_p[:m], out_c[:m]


def create_from_pcd(points: torch.Tensor, colors: torch.Tensor, max_sh_degree: int = 3) -> dict:
    """``CFGaussianModel.create_from_pcd`` (scene/gaussian_model_cf.py:178-212)
    on device-resident float32 points and colours.  The raw parameters are:

      * xyz = points;
      * features [P, (D+1)^2, 3] with the DC row RGB2SH(colour) and the rest 0
        (the module's _features_dc / _features_rest are rows 0 and 1:);
      * scaling = log(sqrt(clamp_min(distCUDA2(points), 1e-7))) repeated x3;
      * rotation (1, 0, 0, 0);
      * opacity inverse_sigmoid(0.1).

    ``dist2`` (distCUDA2's output) is returned as well."""
    device = require_device(points, "points")
    pts = points.contiguous().float()
    col = colors.contiguous().float()
    P = pts.shape[0]
    if col.shape != pts.shape:
        raise ValueError("points and colors must both be [P, 3]")
    K = (int(max_sh_degree) + 1) ** 2
    dist2 = knn3_meandist(pts)
    out = dict(
        xyz=torch.empty((P, 3), dtype=torch.float32, device=device),
        features=torch.empty((P, K, 3), dtype=torch.float32, device=device),
        scaling=torch.empty((P, 3), dtype=torch.float32, device=device),
        rotation=torch.empty((P, 4), dtype=torch.float32, device=device),
        opacity=torch.empty((P, 1), dtype=torch.float32, device=device),
    )
    lib = _lib.load()
    _lib.check(lib.gs_create_from_pcd(P, K, _ptr(pts), _ptr(col), _ptr(dist2), _ptr(out["xyz"]),
                                      _ptr(out["features"]), _ptr(out["scaling"]), _ptr(out["rotation"]),
                                      _ptr(out["opacity"]), _stream(device)), "gs_create_from_pcd")
    out["dist2"] = dist2
    return out


def ingest_frame(depth: torch.Tensor, image: torch.Tensor, intrinsics,
                 voxel_size: Optional[float] = VOXEL_SIZE, max_sh_degree: int = 3,
                 channels_first: Optional[bool] = None) -> dict:
    """One frame to local Gaussians: ``prepare_data``'s lift and voxel
    downsample (trainer/trainer.py:476-512), then ``create_from_pcd``.
    ``voxel_size=None`` keeps every pixel (``down_sample=False``).  Returns the
    raw parameters of ``create_from_pcd`` plus ``points`` and ``colors``."""
    if voxel_size is None:
        pts, cols = depth_to_3d(depth, intrinsics, image, channels_first)
    else:
        pts = depth_to_3d(depth, intrinsics)
        pts, cols = voxel_down_sample(pts, image, voxel_size, channels_first)
    out = create_from_pcd(pts, cols, max_sh_degree)
    out["points"] = pts
    out["colors"] = cols
    return out
