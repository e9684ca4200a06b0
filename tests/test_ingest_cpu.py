"""Frame ingest on the CPU: the oracle pinned against hand-worked answers and
invariants, and the host logic of ``gs_hip.ingest`` (grid sizing, argument
errors, the loud failure on CPU tensors).

The oracle restates kornia's depth_to_3d and open3d's voxel_down_sample.
Both libraries are absent and unpinned (requirements.txt:8-9).  These tests
pin:

  * the lift by reprojection: fx X / Z + cx = u and Z = d, on a non-square
    frame so that a swapped u / v fails;
  * the voxel grid by a hand-worked 1-D case.  Dropping the half-voxel shift
    of the minimum bound, or using round instead of floor, regroups it;
  * the voxel means by a pure-Python grouping (dict, math.floor, math.fsum),
    and by the conservation of the count and the first moment;
  * create_from_pcd by its inverses: SH2RGB(RGB2SH(c)) = c,
    sigmoid(opacity) = 0.1, exp(scale)^2 = clamp(dist2, 1e-7).
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import ingest as OI


def test_oracle_lift_reprojects_to_its_pixel():
    W, H = 7, 5
    rng = np.random.default_rng(0)
    d = rng.uniform(0.5, 4.0, (H, W)).astype(np.float32)
    fx, fy, cx, cy = 6.5, 5.25, 3.1, 2.2
    p = OI.depth_to_3d(d, fx, fy, cx, cy).astype(np.float64)
    assert p.shape == (W * H, 3)
    v, u = np.divmod(np.arange(W * H), W)  # point i = pixel (u, v), i = v * W + u
    np.testing.assert_array_equal(p[:, 2], d.reshape(-1))
    np.testing.assert_allclose(fx * p[:, 0] / p[:, 2] + cx, u, rtol=0, atol=1e-5)
    np.testing.assert_allclose(fy * p[:, 1] / p[:, 2] + cy, v, rtol=0, atol=1e-5)
    # kornia's operation order, bit for bit: ((u - cx) / fx) * d in float32
    i = 3 * W + 5
    x = (np.float32(5) - np.float32(cx)) / np.float32(fx)
    assert np.float32(p[i, 0]) == x * d[3, 5]


def test_oracle_voxel_hand_worked_grid():
    # 1-D case along x with s = 0.01: min 0 -> voxel_min = -0.005, so
    # ref = (x + 0.005) / 0.01 = 0.5, 0.9, 1.1, 1.7, 1.99, 2.01 and
    # the voxels are {0, 0.004}, {0.006, 0.012, 0.0149}, {0.0151}
    xs = np.array([0.0, 0.004, 0.006, 0.012, 0.0149, 0.0151], np.float32)
    pts = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], 1)
    cols = np.arange(18, dtype=np.float64).reshape(6, 3) / 17.0
    P, C, idx, cnt = OI.voxel_down_sample(pts, cols, 0.01)
    assert idx[:, 0].tolist() == [0, 1, 2] and cnt.tolist() == [2, 3, 1]
    exp = [np.float64(xs[0]) + np.float64(xs[1]), np.float64(xs[2]) + np.float64(xs[3]) + np.float64(xs[4]),
           np.float64(xs[5])]
    np.testing.assert_array_equal(P[:, 0], np.array([exp[0] / 2, exp[1] / 3, exp[2]], np.float64).astype(np.float32))
    np.testing.assert_array_equal(P[:, 1:], 0.0)
    np.testing.assert_allclose(C[1], cols[2:5].mean(0), rtol=1e-7)


def test_oracle_voxel_matches_pure_python_grouping():
    rng = np.random.default_rng(3)
    N, s = 400, 0.25
    pts = rng.normal(0.0, 0.6, (N, 3)).astype(np.float32)
    cols = rng.integers(0, 256, (N, 3)).astype(np.uint8)
    P, C, idx, cnt = OI.voxel_down_sample(pts, OI.colors_u8(cols), s)
    mn = [min(float(p[a]) for p in pts) for a in range(3)]
    groups = {}
    for i, p in enumerate(pts):
        key = tuple(int(math.floor((float(p[a]) - (mn[a] - s * 0.5)) / s)) for a in range(3))
        groups.setdefault(key, []).append(i)
    keys = sorted(groups)
    assert [tuple(k) for k in idx.tolist()] == keys
    for m, k in enumerate(keys):
        members = groups[k]
        assert cnt[m] == len(members)
        for a in range(3):
            mean_p = math.fsum(float(pts[i, a]) for i in members) / len(members)
            mean_c = math.fsum(int(cols[i, a]) / 255.0 for i in members) / len(members)
            assert abs(float(P[m, a]) - mean_p) <= 1e-7 * max(1.0, abs(mean_p))
            assert abs(float(C[m, a]) - mean_c) <= 1e-7


def test_oracle_voxel_invariants():
    rng = np.random.default_rng(5)
    N, s = 20_000, 0.05
    pts = (rng.normal(0.0, 1.0, (N, 3)) * [1.0, 0.5, 2.0]).astype(np.float32)
    cols = rng.random((N, 3))
    P, C, idx, cnt = OI.voxel_down_sample(pts, cols, s)
    assert cnt.sum() == N and np.all(cnt >= 1)
    # ascending (ix, iy, iz) and unique
    flat = (idx[:, 0] * 2 ** 40 + idx[:, 1] * 2 ** 20 + idx[:, 2])
    assert np.all(np.diff(flat) > 0)
    # the count-weighted means give back the first moment
    np.testing.assert_allclose((P.astype(np.float64) * cnt[:, None]).sum(0), pts.astype(np.float64).sum(0),
                               rtol=0, atol=1e-3)
    # each mean lies in its voxel (closed box, float32 rounding of the mean)
    vmin = pts.astype(np.float64).min(0) - s * 0.5
    lo = vmin + idx * s
    assert np.all(P >= lo - 1e-6) and np.all(P <= lo + s + 1e-6)
    # one voxel for everything; one voxel per point
    P1, _, _, c1 = OI.voxel_down_sample(pts, cols, 1e4)
    assert c1.tolist() == [N]
    P2, _, _, c2 = OI.voxel_down_sample(pts[:50], cols[:50], 1e-6)
    assert c2.tolist() == [1] * 50 and sorted(map(tuple, P2.tolist())) == sorted(map(tuple, pts[:50].tolist()))


def test_oracle_voxel_errors():
    pts = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    with pytest.raises(ValueError, match="voxel_size <= 0"):
        OI.voxel_down_sample(pts, np.zeros((2, 3)), 0.0)
    with pytest.raises(ValueError, match="too small"):
        OI.voxel_down_sample(pts, np.zeros((2, 3)), 1e-12)


def test_oracle_create_from_pcd_inverts():
    rng = np.random.default_rng(7)
    P = 1000
    pts = rng.normal(size=(P, 3)).astype(np.float32)
    col = rng.random((P, 3)).astype(np.float32)
    d2 = rng.uniform(0.0, 1e-2, P).astype(np.float32)
    d2[:10] = 1e-9  # clamped to 1e-7
    out = OI.create_from_pcd(pts, col, d2, 3)
    f = out["features"]
    assert f.shape == (P, 16, 3) and np.all(f[:, 1:] == 0)
    np.testing.assert_allclose(f[:, 0].astype(np.float64) * OI.SH_C0 + 0.5, col, rtol=0, atol=1e-6)
    sig = 1.0 / (1.0 + math.exp(-float(out["opacity"][0, 0])))
    assert abs(sig - 0.1) < 1e-7 and np.all(out["opacity"] == out["opacity"][0, 0])
    np.testing.assert_allclose(np.exp(2.0 * out["scaling"].astype(np.float64)),
                               np.repeat(np.maximum(d2, np.float32(1e-7))[:, None], 3, 1), rtol=2e-6)
    assert np.all(out["rotation"] == np.array([1, 0, 0, 0], np.float32))
    np.testing.assert_array_equal(out["xyz"], pts)
    assert OI.create_from_pcd(pts, col, d2, 0)["features"].shape == (P, 1, 3)


def test_voxel_grid_host_logic():
    from gs_hip.ingest import voxel_grid
    vmin, bits = voxel_grid([0.0, 0.0, 0.0], [0.0149, 1.0, 0.0], 0.01)
    np.testing.assert_array_equal(vmin, [-0.005, -0.005, -0.005])
    # top indices floor(1.99) = 1, floor(100.5) = 100, floor(0.5) = 0
    assert bits == (1, 7, 0)
    with pytest.raises(ValueError, match="too small"):
        voxel_grid([0, 0, 0], [1, 1, 1], 1e-12)
    with pytest.raises(ValueError, match="key bits"):
        voxel_grid([0, 0, 0], [1, 1, 1], 1e-9)  # 30 bits per axis
    with pytest.raises(ValueError, match="<= 0"):
        voxel_grid([0, 0, 0], [1, 1, 1], -1.0)
    # the host grid holds every index the oracle produces
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(5000, 3)).astype(np.float32) * 3
    vmin, bits = voxel_grid(pts.min(0).astype(np.float64), pts.max(0).astype(np.float64), 0.02)
    _, _, idx, _ = OI.voxel_down_sample(pts, np.zeros((5000, 3)), 0.02)
    assert all(int(idx[:, a].max()) < 2 ** bits[a] for a in range(3))


def test_ingest_c_abi_rejects_bad_arguments_without_device():
    from gs_hip import _lib
    lib = _lib.load()
    vm = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    assert lib.gs_voxel_down_sample(None, 0, None, 0, 3, 1, vm, 0.01, 1, 1, 1, None, None, None, None, 0, None) == 1
    assert lib.gs_create_from_pcd(-1, 16, None, None, None, None, None, None, None, None, None) == 1
    assert lib.gs_create_from_pcd(5, 17, None, None, None, None, None, None, None, None, None) == 1
    assert lib.gs_depth_to_points(4, 4, None, 1.0, 1.0, 0.0, 0.0, None, None, 0, 3, 1, None, None) == 1
    assert lib.gs_points_bounds(None, 0, None, None, 0, None) == 1
    assert lib.gs_voxel_workspace_bytes(1000) > 1000 * 24


def test_ingest_cpu_tensors_fail_loudly():
    from gs_hip import ingest
    with pytest.raises(RuntimeError, match="HIP device"):
        ingest.depth_to_3d(torch.ones(4, 4), (1.0, 1.0, 2.0, 2.0))
    with pytest.raises(RuntimeError, match="HIP device"):
        ingest.voxel_down_sample(torch.ones(4, 3), torch.ones(4, 3))
