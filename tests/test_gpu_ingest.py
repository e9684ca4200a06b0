"""Frame ingest on the GPU against the CPU oracle (oracle/ingest.py).

  * depth lift: bit-exact points, and colours from 8-bit or float images;
  * bounds: exact min / max, and a non-finite cloud is refused;
  * voxel downsample: bit-exact means and colours and the same voxel count,
    on lifted frames, on random clouds, and at the extremes (one voxel per
    point, one voxel for all, one point, a grid wider than 32 key bits);
  * create_from_pcd: bit for bit the torch ops of
    scene/gaussian_model_cf.py:178-212 run on the GPU, and the oracle within
    1 ulp of the float32 log;
  * the whole frame: ingest -> render gives a finite image, and two calls
    agree bit for bit.
"""
import math

import numpy as np
import pytest
import torch

from gs_hip import ingest, raster
from gs_hip.camera import focal2fov, fov2focal, make_camera, raster_settings
from gs_hip.scenes import inverse_sigmoid, rgb2sh, synthetic_depth
from oracle import ingest as OI
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def frame(seed, W, H):
    g = torch.Generator().manual_seed(seed)
    d = synthetic_depth(W, H, g)
    img = torch.randint(0, 256, (H, W, 3), generator=g, dtype=torch.uint8)
    fx = fov2focal(math.radians(60.0), W)
    return d, img, (fx, fx, W / 2.0, H / 2.0)


def f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("W,H", [(480, 270), (7, 5), (1, 1)])
def test_lift_bit_exact(W, H):
    d, img, (fx, fy, cx, cy) = frame(W * 31 + H, W, H)
    pts, cols = ingest.depth_to_3d(d.to(DEV), (fx, fy, cx, cy), img.to(DEV))
    ref = OI.depth_to_3d(d.numpy(), f32(fx), f32(fy), f32(cx), f32(cy))
    np.testing.assert_array_equal(pts.cpu().numpy(), ref)
    np.testing.assert_array_equal(cols.cpu().numpy(), OI.colors_u8(img.numpy()).astype(np.float32))
    # a float32 [3, H, W] image (the torch layout of Camera.original_image)
    chw = (img.permute(2, 0, 1).float() / 255.0).contiguous()
    _, cols2 = ingest.depth_to_3d(d.to(DEV), (fx, fy, cx, cy), chw.to(DEV))
    np.testing.assert_array_equal(cols2.cpu().numpy(), chw.permute(1, 2, 0).reshape(-1, 3).numpy())
    # a 3x3 intrinsics matrix reads the same four entries
    K = torch.tensor([[fx, 0.0, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]])
    torch.testing.assert_close(ingest.depth_to_3d(d.to(DEV), K), pts, rtol=0, atol=0)


def test_bounds_exact_and_non_finite_refused():
    rng = np.random.default_rng(1)
    for n in (1, 255, 256, 70_001, 600_000):
        p = rng.normal(size=(n, 3)).astype(np.float32) * [1.0, 10.0, 0.1]
        p = p.astype(np.float32)
        mn, mx = ingest.points_bounds(torch.from_numpy(p).to(DEV))
        np.testing.assert_array_equal(mn, p.min(0).astype(np.float64))
        np.testing.assert_array_equal(mx, p.max(0).astype(np.float64))
    p = rng.normal(size=(1000, 3)).astype(np.float32)
    p[777, 1] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        ingest.points_bounds(torch.from_numpy(p).to(DEV))
    p[777, 1] = np.inf
    with pytest.raises(ValueError, match="non-finite"):
        ingest.points_bounds(torch.from_numpy(p).to(DEV))


def check_voxels(pts, cols_dev, cols_ref64, s, channels_first=None):
    out_p, out_c = ingest.voxel_down_sample(pts.to(DEV), cols_dev, s, channels_first)
    P, C, idx, cnt = OI.voxel_down_sample(pts.cpu().numpy(), cols_ref64, s)
    assert out_p.shape[0] == P.shape[0], (out_p.shape, P.shape)
    np.testing.assert_array_equal(out_p.cpu().numpy(), P)
    np.testing.assert_array_equal(out_c.cpu().numpy(), C)
    return out_p, out_c


@pytest.mark.parametrize("W,H,s", [(480, 270, 0.01), (960, 540, 0.01), (480, 270, 0.05), (64, 48, 1e4)])
def test_voxel_down_sample_frame_bit_exact(W, H, s):
    d, img, K = frame(W + H, W, H)
    pts = ingest.depth_to_3d(d.to(DEV), K)
    check_voxels(pts, img.to(DEV), OI.colors_u8(img.numpy()), s)
    chw = (img.permute(2, 0, 1).float() / 255.0).contiguous()
    ref = chw.permute(1, 2, 0).reshape(-1, 3).double().numpy()
    check_voxels(pts, chw.to(DEV), ref, s)


@pytest.mark.parametrize("case", ["normal", "own_voxel", "one_point", "wide_grid", "negative"])
def test_voxel_down_sample_clouds_bit_exact(case):
    rng = np.random.default_rng(["normal", "own_voxel", "one_point", "wide_grid", "negative"].index(case))
    if case == "normal":
        p, s = rng.normal(size=(100_000, 3)), 0.05
    elif case == "own_voxel":
        p, s = rng.uniform(-1, 1, (20_000, 3)), 1e-6
    elif case == "one_point":
        p, s = np.array([[0.25, -3.0, 7.5]]), 0.01
    elif case == "wide_grid":  # 20 bits per axis -> 60-bit keys
        p, s = rng.uniform(-500, 500, (50_000, 3)), 1e-3
    else:
        p, s = -np.abs(rng.normal(size=(30_000, 3))) - 5.0, 0.02
    pts = torch.from_numpy(p.astype(np.float32))
    cols = rng.random((pts.shape[0], 3)).astype(np.float32)
    out_p, _ = check_voxels(pts, torch.from_numpy(cols).to(DEV), cols.astype(np.float64), s)
    if case == "wide_grid":
        mn, mx = ingest.points_bounds(pts.to(DEV))
        assert sum(ingest.voxel_grid(mn, mx, s)[1]) > 32
    if case == "own_voxel":
        assert out_p.shape[0] == pts.shape[0]


def test_create_from_pcd_matches_torch_ops():
    d, img, K = frame(5, 480, 270)
    pts, cols = ingest.voxel_down_sample(ingest.depth_to_3d(d.to(DEV), K), img.to(DEV), 0.01)
    out = ingest.create_from_pcd(pts, cols, 3)
    P = pts.shape[0]
    # the reference's own torch ops on the device (scene/gaussian_model_cf.py:182-203)
    dist2 = torch.clamp_min(raster.knn3_meandist(pts), 0.0000001)
    scales = torch.log(torch.sqrt(dist2))[..., None].repeat(1, 3)
    fused_color = rgb2sh(cols)
    feats = torch.zeros((P, 3, 16), device=DEV)
    feats[:, :3, 0] = fused_color[:, :3]
    opac = inverse_sigmoid(0.1 * torch.ones((P, 1), dtype=torch.float, device=DEV))
    bad = {k: int((a != b).sum()) for k, a, b in (
        ("xyz", out["xyz"], pts),
        ("f_dc", out["features"][:, :1], feats[:, :, 0:1].transpose(1, 2)),
        ("f_rest", out["features"][:, 1:], feats[:, :, 1:].transpose(1, 2)),
        ("scaling", out["scaling"], scales),
        ("opacity", out["opacity"], opac),
        ("rotation", out["rotation"], torch.tensor([1.0, 0, 0, 0], device=DEV).expand(P, 4)))}
    assert not any(bad.values()), bad
    # the oracle: exact except the float32 log (1 ulp)
    ref = OI.create_from_pcd(pts.cpu().numpy(), cols.cpu().numpy(), out["dist2"].cpu().numpy(), 3)
    np.testing.assert_array_equal(out["features"].cpu().numpy(), ref["features"])
    np.testing.assert_allclose(out["scaling"].cpu().numpy(), ref["scaling"], rtol=2.5e-7, atol=0)
    np.testing.assert_allclose(out["opacity"].cpu().numpy(), ref["opacity"], rtol=2.5e-7, atol=0)
    assert ingest.create_from_pcd(pts, cols, 0)["features"].shape == (P, 1, 3)


def test_create_from_pcd_dist2_is_the_brute_force_knn():
    rng = np.random.default_rng(9)
    p = rng.normal(size=(3000, 3)).astype(np.float32)
    out = ingest.create_from_pcd(torch.from_numpy(p).to(DEV), torch.rand(3000, 3).to(DEV), 3)
    np.testing.assert_array_equal(out["dist2"].cpu().numpy(), O.knn3_brute(p))


def test_ingest_frame_renders_and_is_deterministic():
    W, H = 480, 270
    d, img, K = frame(21, W, H)
    a = ingest.ingest_frame(d.to(DEV), img.to(DEV), K, voxel_size=0.01)
    b = ingest.ingest_frame(d.to(DEV), img.to(DEV), K, voxel_size=0.01)
    for k in ("xyz", "features", "scaling", "rotation", "opacity"):
        assert torch.equal(a[k], b[k]), k
    _, _, _, cnt = OI.voxel_down_sample(OI.depth_to_3d(d.numpy(), *[f32(x) for x in K]), OI.colors_u8(img.numpy()),
                                        0.01)
    assert a["xyz"].shape[0] == cnt.shape[0]
    full = ingest.ingest_frame(d.to(DEV), img.to(DEV), K, voxel_size=None)
    assert full["xyz"].shape[0] == W * H
    fovx = math.radians(60.0)
    cam = make_camera(np.eye(3), np.zeros(3), fovx, focal2fov(K[1], H), W, H).to(DEV)
    rs = raster_settings(cam, torch.zeros(3, device=DEV), 0)
    for g in (a, full):
        color, radii, depth, _, _ = raster.forward(rs, g["xyz"], g["features"][:, :1].contiguous(), None,
                                                   torch.sigmoid(g["opacity"]), torch.exp(g["scaling"]),
                                                   g["rotation"], None, want_alpha=False)
        assert torch.isfinite(color).all() and torch.isfinite(depth).all()
        assert (radii > 0).sum() > 0.9 * g["xyz"].shape[0]
