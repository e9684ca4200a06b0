"""CPU oracle for frame ingest -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of how the reference turns a depth map and its image
into the initial Gaussians of a frame.  Only ``tests/`` and the CPU-baseline
leg of ``bench.py`` may import it, as the checker or the timed CPU baseline;
the product path (``gs_hip.ingest``) never routes through it.

The steps follow these sources:

  * the lift follows kornia ``depth_to_3d`` as trainer/trainer.py:489-495
    calls it (``normalize_points=False``).  kornia is an unpinned dependency
    (requirements.txt:8) and is absent here.  Its published
    ``unproject_points`` computes x = (u - cx) / fx, y = (v - cy) / fy in the
    depth's dtype and multiplies [x, y, 1] by the depth.  The pixel grid is
    ``create_meshgrid(normalized_coordinates=False)``: u = 0 .. W-1 along a
    row, v = 0 .. H-1 down the columns.
  * the voxel downsample follows open3d ``PointCloud::VoxelDownSample`` as
    trainer/trainer.py:414-415,506-507 calls it (voxel 0.01).  open3d is an
    unpinned dependency (requirements.txt:9) and is absent here.  Its
    published algorithm runs in double precision:
      - voxel_min_bound = min_bound - s/2;
      - it fails with "voxel_size is too small" when s * INT_MAX < the largest
        extent of [min - s/2, max + s/2];
      - index = int(floor((p - voxel_min_bound) / s));
      - each voxel's mean point and mean colour are double sums in point
        order (AccumulatedPoint::AddPoint) divided by the count.
    The caller then converts with ``np.asarray(..., dtype=np.float32)``.
    Voxels are returned here in ascending (ix, iy, iz) order; open3d's order
    is that of its hash map, which is unspecified.
  * the colours follow ``np.asarray(image) / 255.0`` (trainer/trainer.py:401,450).
  * ``create_from_pcd`` follows scene/gaussian_model_cf.py:178-212, as the
    torch ops compute it on the device:
      - RGB2SH = (c - 0.5) / C0 (utils/sh_utils.py:114-115).  torch's
        division of a float32 tensor by a host scalar multiplies by the
        reciprocal, taken in double and rounded to float32.  Measured on the
        GPU: this matches torch bit for bit, while a float32 reciprocal,
        a float32 division and a double division each differ on 38-96% of
        the elements;
      - scale = log(sqrt(clamp_min(dist2, 1e-7))) repeated x3;
      - rotation (1, 0, 0, 0);
      - opacity = inverse_sigmoid(0.1) = log(0.1 / (1 - 0.1)) in float32
        (utils/general_utils.py:16-17).
"""
from __future__ import annotations

import numpy as np

INT_MAX = 2 ** 31 - 1
SH_C0 = 0.28209479177387814


def depth_to_3d(depth, fx, fy, cx, cy) -> np.ndarray:
    """[H, W] float32 depth -> [H*W, 3] float32 points (row-major pixels)."""
    d = np.asarray(depth, dtype=np.float32)
    H, W = d.shape
    f32 = np.float32
    u = np.arange(W, dtype=np.float32)
    v = np.arange(H, dtype=np.float32)
    x = (u - f32(cx)) / f32(fx)
    y = (v - f32(cy)) / f32(fy)
    X = x[None, :] * d
    Y = y[:, None] * d
    return np.stack([X, Y, d], axis=-1).reshape(-1, 3).astype(np.float32)


def colors_u8(image) -> np.ndarray:
    """np.asarray(image) / 255.0 for an 8-bit [.., 3] image: float64 [N, 3]."""
    return (np.asarray(image, dtype=np.uint8).reshape(-1, 3) / 255.0).astype(np.float64)


def voxel_down_sample(points, colors, voxel_size: float):
    """Returns (points f32 [M, 3], colors f32 [M, 3], voxel index int64 [M, 3],
    counts int64 [M]); colours are taken as given, in float64."""
    p = np.asarray(points, dtype=np.float32).astype(np.float64).reshape(-1, 3)
    c = np.asarray(colors, dtype=np.float64).reshape(-1, 3)
    s = float(voxel_size)
    if not s > 0.0:
        raise ValueError("voxel_size <= 0.")
    mn = p.min(axis=0)
    mx = p.max(axis=0)
    vmin = mn - s * 0.5
    vmax = mx + s * 0.5
    if s * INT_MAX < float(np.max(vmax - vmin)):
        raise ValueError("voxel_size is too small.")
    idx = np.floor((p - vmin) / s).astype(np.int64)
    uniq, inv = np.unique(idx, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    M = uniq.shape[0]
    sp = np.zeros((M, 3), dtype=np.float64)
    sc = np.zeros((M, 3), dtype=np.float64)
    np.add.at(sp, inv, p)  # unbuffered, in point order
    np.add.at(sc, inv, c)
    cnt = np.bincount(inv, minlength=M).astype(np.int64)
    return ((sp / cnt[:, None]).astype(np.float32), (sc / cnt[:, None]).astype(np.float32), uniq, cnt)


def create_from_pcd(points, colors, dist2, max_sh_degree: int = 3) -> dict:
    """Raw parameters of create_from_pcd from points, colours and distCUDA2."""
    f32 = np.float32
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    col = np.asarray(colors, dtype=np.float32).reshape(-1, 3)
    d2 = np.asarray(dist2, dtype=np.float32).reshape(-1)
    P = pts.shape[0]
    K = (max_sh_degree + 1) ** 2
    inv_c0 = f32(1.0 / SH_C0)
    feats = np.zeros((P, K, 3), dtype=np.float32)
    feats[:, 0, :] = (col - f32(0.5)) * inv_c0
    with np.errstate(invalid="ignore"):
        s = np.log(np.sqrt(np.maximum(d2, f32(1e-7))))
    x = f32(0.1)
    opac = np.log(x / (f32(1.0) - x))
    rot = np.zeros((P, 4), dtype=np.float32)
    rot[:, 0] = 1.0
    return dict(xyz=pts.copy(), features=feats, scaling=np.repeat(s[:, None], 3, axis=1).astype(np.float32),
                rotation=rot, opacity=np.full((P, 1), opac, dtype=np.float32))
