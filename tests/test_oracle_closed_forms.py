"""Pins the CPU oracle's forward against closed forms and independent
libraries (scipy SH / rotations / kd-tree, numpy stable sort, a numpy
brute-force compositor), so that a dropped term, a wrong sign or index, or a
transposed operand in the restatement fails here.
"""
import math

import numpy as np
import pytest
from scipy import special
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from gs_hip.camera import make_camera, oracle_camera
from oracle import oracle as O


def cam_dict(W=64, H=48, fovx_deg=60.0, deg=0, bg=(0.0, 0.0, 0.0)):
    fovx = math.radians(fovx_deg)
    fovy = 2 * math.atan(math.tan(fovx / 2) * H / W)
    cam = make_camera(np.eye(3), np.zeros(3), fovx, fovy, W, H)
    return oracle_camera(cam, np.asarray(bg, np.float32), deg)


def one(means, scales, opac=0.5, rot=(1, 0, 0, 0), colors=None):
    means = np.atleast_2d(np.asarray(means, np.float64))
    P = means.shape[0]
    sc = dict(means=means, scales=np.broadcast_to(np.asarray(scales, np.float64), (P, 3)).copy(),
              rotations=np.broadcast_to(np.asarray(rot, np.float64), (P, 4)).copy(),
              opacities=np.broadcast_to(np.asarray(opac, np.float64), (P,)).copy())
    sc["colors"] = np.full((P, 3), 0.5) if colors is None else np.asarray(colors, np.float64)
    return sc


# ---------------------------------------------------------------- exp
def test_spec_expf_accuracy():
    x = np.concatenate([np.linspace(-80, 0, 200001, dtype=np.float32), np.float32([0.0, -0.0, -1e-30])])
    y = O.spec_expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.abs(y - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 4.0, ulp.max()
    assert O.spec_expf(np.float32([0.0]))[0] == 1.0


# ---------------------------------------------------------------- projection / EWA
def test_on_axis_isotropic_radius_and_center():
    cam = cam_dict()
    W, H = cam["W"], cam["H"]
    z, s = 4.0, 0.2
    sc = one([0.0, 0.0, z], [s, s, s])
    g = O.preprocess(sc, cam, np.float64)
    f = W / (2 * cam["tanfovx"])
    lam = f * f * s * s / (z * z) + 0.3
    # mid^2 - det = 0 -> floor sqrt(0.1)
    r = math.ceil(3 * math.sqrt(lam + math.sqrt(0.1)))
    assert g["radii"][0] == r
    np.testing.assert_allclose(g["xy"][0], [(W - 1) / 2, (H - 1) / 2], atol=1e-9)
    np.testing.assert_allclose(g["conic_op"][0, :3], [1 / lam, 0.0, 1 / lam], rtol=1e-9, atol=1e-12)
    assert g["gdepth"][0] == pytest.approx(z)


def test_pinhole_projection_and_ewa_covariance():
    cam = cam_dict(W=80, H=60)
    W, H = cam["W"], cam["H"]
    fx = W / (2 * cam["tanfovx"])
    fy = H / (2 * cam["tanfovy"])
    rng = np.random.default_rng(0)
    P = 50
    means = np.stack([rng.uniform(-1, 1, P), rng.uniform(-0.7, 0.7, P), rng.uniform(3, 6, P)], 1)
    scales = rng.uniform(0.05, 0.3, (P, 3))
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    sc = dict(means=means, scales=scales, rotations=q, opacities=np.full(P, 0.5), colors=np.full((P, 3), 0.5))
    g = O.preprocess(sc, cam, np.float64)
    x, y, z = means.T
    # the projection matrix is float32 (as the reference's tensors are): ~1e-7 relative
    np.testing.assert_allclose(g["xy"][:, 0], fx * x / z + (W - 1) / 2, rtol=0, atol=1e-5)
    np.testing.assert_allclose(g["xy"][:, 1], fy * y / z + (H - 1) / 2, rtol=0, atol=1e-5)
    np.testing.assert_allclose(g["gdepth"], z, rtol=1e-12)
    # Sigma = R S S R^T with R from scipy (scalar-last quaternion), J the pinhole Jacobian
    Rm = Rotation.from_quat(np.concatenate([q[:, 1:], q[:, :1]], 1)).as_matrix()
    Sig = Rm @ np.einsum("pi,ij->pij", scales ** 2, np.eye(3)) @ Rm.transpose(0, 2, 1)
    J = np.zeros((P, 2, 3))
    J[:, 0, 0] = fx / z
    J[:, 0, 2] = -fx * x / z ** 2
    J[:, 1, 1] = fy / z
    J[:, 1, 2] = -fy * y / z ** 2
    C = J @ Sig @ J.transpose(0, 2, 1) + 0.3 * np.eye(2)
    Ci = np.linalg.inv(C)
    np.testing.assert_allclose(g["conic_op"][:, 0], Ci[:, 0, 0], rtol=1e-9)
    np.testing.assert_allclose(g["conic_op"][:, 1], Ci[:, 0, 1], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g["conic_op"][:, 2], Ci[:, 1, 1], rtol=1e-9)
    cov6 = np.stack([Sig[:, 0, 0], Sig[:, 0, 1], Sig[:, 0, 2], Sig[:, 1, 1], Sig[:, 1, 2], Sig[:, 2, 2]], 1)
    np.testing.assert_allclose(g["cov3D"], cov6, rtol=1e-10, atol=1e-14)
    lam_max = np.linalg.eigvalsh(C).max(axis=1)
    disc = np.sqrt(np.maximum(0.1, ((C[:, 0, 0] + C[:, 1, 1]) / 2) ** 2 - np.linalg.det(C)))
    r_ref = np.ceil(3 * np.sqrt((C[:, 0, 0] + C[:, 1, 1]) / 2 + disc))
    np.testing.assert_array_equal(g["radii"], r_ref.astype(np.int32))
    assert np.all(r_ref >= np.ceil(3 * np.sqrt(lam_max)))


def test_frustum_clamp_and_near_cull():
    cam = cam_dict()
    sc = one([[0, 0, 0.1], [0, 0, 0.2], [0, 0, 0.2001], [50.0, 0, 4.0]], [0.1, 0.1, 0.1])
    g = O.preprocess(sc, cam, np.float64)
    assert g["radii"][0] == 0 and g["radii"][1] == 0 and g["radii"][2] > 0
    # far outside the image horizontally: its rect is empty -> culled
    assert g["radii"][3] == 0 and g["tiles"][3] == 0


def test_pose_equals_pretransformed_means():
    cam = cam_dict(W=64, H=64, deg=0)
    rng = np.random.default_rng(2)
    P = 60
    means = np.stack([rng.uniform(-0.8, 0.8, P), rng.uniform(-0.8, 0.8, P), rng.uniform(3, 5, P)], 1)
    sc = one(means, [0.1, 0.1, 0.1], colors=rng.uniform(0, 1, (P, 3)))
    qxyzw = np.array([0.05, -0.1, 0.03, 0.99])
    t = np.array([0.1, -0.05, 0.2])
    cam["pose7"] = np.concatenate([t, qxyzw * 1.7])  # unnormalised on purpose
    a = O.forward(sc, cam, np.float64)
    R = Rotation.from_quat(qxyzw / np.linalg.norm(qxyzw)).as_matrix()
    sc2 = dict(sc)
    sc2["means"] = means @ R.T + t
    cam2 = dict(cam)
    del cam2["pose7"]
    b = O.forward(sc2, cam2, np.float64)
    np.testing.assert_allclose(a["color"], b["color"], atol=1e-10)
    np.testing.assert_array_equal(a["radii"], b["radii"])


# ---------------------------------------------------------------- SH
def real_sh_scipy(l, m, d):
    """Real SH in the reference's basis: m>0 sqrt2 Re Y_l^m, m<0 sqrt2 Im Y_l^|m|,
    m=0 Y_l^0, with scipy's (Condon-Shortley) complex harmonics."""
    polar = np.arccos(np.clip(d[:, 2], -1, 1))
    azim = np.arctan2(d[:, 1], d[:, 0])
    Y = special.sph_harm_y(l, abs(m), polar, azim)
    if m > 0:
        return math.sqrt(2) * Y.real
    if m < 0:
        return math.sqrt(2) * Y.imag
    return Y.real


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_basis_matches_scipy(deg):
    cam = cam_dict(W=64, H=64, deg=deg)
    rng = np.random.default_rng(deg)
    P = 40
    means = np.stack([rng.uniform(-1, 1, P), rng.uniform(-1, 1, P), rng.uniform(3, 5, P)], 1)
    d = means - cam["campos"]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for l in range(deg + 1):
        for m in range(-l, l + 1):
            k = l * l + l + m
            shs = np.zeros((P, 16, 3))
            shs[:, k, 0] = 0.1
            sc = dict(means=means, scales=np.full((P, 3), 0.05), rotations=np.tile([1.0, 0, 0, 0], (P, 1)),
                      opacities=np.full(P, 0.5), shs=shs)
            g = O.preprocess(sc, cam, np.float64)
            vis = g["radii"] > 0
            assert vis.sum() > 10
            np.testing.assert_allclose(g["rgb"][vis, 0] - 0.5, 0.1 * real_sh_scipy(l, m, d)[vis], atol=1e-12,
                                       err_msg=f"l={l} m={m}")
            np.testing.assert_allclose(g["rgb"][vis, 1:], 0.5, atol=1e-15)


def test_sh_clamp_mask():
    cam = cam_dict(W=64, H=64, deg=0)
    shs = np.zeros((2, 16, 3))
    shs[0, 0] = [-10.0, 0.0, 10.0]
    shs[1, 0] = [1.0, -2.0, 0.0]  # C0 * -2 + 0.5 < 0
    sc = dict(means=np.array([[0, 0, 4.0], [0.2, 0, 4.0]]), scales=np.full((2, 3), 0.1),
              rotations=np.tile([1.0, 0, 0, 0], (2, 1)), opacities=np.full(2, 0.5), shs=shs)
    g = O.preprocess(sc, cam, np.float64)
    np.testing.assert_array_equal(g["clamped"], [[1, 0, 0], [0, 1, 0]])
    assert g["rgb"][0, 0] == 0.0 and g["rgb"][1, 1] == 0.0


# ---------------------------------------------------------------- compositing
def test_single_and_two_gaussian_compositing():
    cam = cam_dict(W=48, H=48, bg=(0.2, 0.4, 0.6))
    sc = one([[0, 0, 4.0], [0.05, 0, 5.0]], [0.15, 0.15, 0.15], opac=[0.7, 0.6],
             colors=[[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    f = O.forward(sc, cam, np.float64)
    H, W = cam["H"], cam["W"]
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)

    def alpha(i):
        d = np.stack([f["xy"][i, 0] - xs, f["xy"][i, 1] - ys])
        a, b, c, o = f["conic_op"][i]
        pw = -0.5 * (a * d[0] ** 2 + c * d[1] ** 2) - b * d[0] * d[1]
        al = np.minimum(0.99, o * np.exp(pw))
        return np.where((pw <= 0) & (al >= 1 / 255), al, 0.0)

    a1, a2 = alpha(0), alpha(1)  # Gaussian 0 is nearer
    bg = cam["bg"].astype(np.float64)
    col = sc["colors"]
    exp_c = (a1[None] * col[0][:, None, None] + (1 - a1)[None] * a2[None] * col[1][:, None, None]
             + ((1 - a1) * (1 - a2))[None] * bg[:, None, None])
    np.testing.assert_allclose(f["color"], exp_c, atol=1e-12)
    np.testing.assert_allclose(f["alpha"][0], 1 - (1 - a1) * (1 - a2), atol=1e-12)
    np.testing.assert_allclose(f["depth"][0], a1 * 4.0 + (1 - a1) * a2 * 5.0, atol=1e-12)


def test_saturation_stops_after_two_opaque_layers():
    """Three 0.99-alpha layers at the centre pixel.  In float32 1 - 0.99 is
    0.00999999, so after the first layer T = 0.00999999 and the second would
    give T*(1-a) = 9.99998e-5 < 1e-4: the pixel ends after one contributor."""
    cam = cam_dict(W=16, H=16)
    c = [(16 - 1) / 2] * 2
    fx = 16 / (2 * cam["tanfovx"])
    # centre each Gaussian on pixel (7, 7): x/z * fx + 7.5 = 7; opacity 1 makes
    # alpha = min(0.99, exp(power)) = 0.99 there
    zs = (4.0, 4.5, 5.0)
    sc = one([[-0.5 * z / fx, -0.5 * z / fx, z] for z in zs], [0.3, 0.3, 0.3], opac=1.0)
    f = O.forward(sc, cam, np.float32)
    T = np.float32(1.0)
    n = 0
    for _ in range(3):
        t2 = np.float32(T * np.float32(np.float32(1.0) - np.float32(0.99)))
        if t2 < np.float32(1e-4):
            break
        T = t2
        n += 1
    assert f["n_contrib"][7, 7] == n == 1
    assert f["final_T"][7, 7] == T


def _np_rect(xy, r, gx, gy):
    x0 = np.clip(np.trunc((xy[:, 0] - r) / np.float32(16)), 0, gx).astype(np.int64)
    y0 = np.clip(np.trunc((xy[:, 1] - r) / np.float32(16)), 0, gy).astype(np.int64)
    x1 = np.clip(np.trunc((((xy[:, 0] + r) + np.float32(16)) - np.float32(1)) / np.float32(16)), 0, gx).astype(np.int64)
    y1 = np.clip(np.trunc((((xy[:, 1] + r) + np.float32(16)) - np.float32(1)) / np.float32(16)), 0, gy).astype(np.int64)
    return x0, y0, x1, y1


def test_binning_and_blend_against_numpy_brute_force():
    from gs_hip import scenes
    sc = scenes.plain_scene(seed=0, P=600, W=70, H=50)
    cam = oracle_camera(sc["cam"], sc["bg"], 3)
    scene = {k: sc[k].numpy() for k in ("means", "scales", "rotations", "opacities", "shs")}
    f = O.forward(scene, cam, np.float32)
    W, H = cam["W"], cam["H"]
    gx, gy = (W + 15) // 16, (H + 15) // 16
    vis = np.nonzero(f["radii"] > 0)[0]
    r = f["radii"].astype(np.float32)
    x0, y0, x1, y1 = _np_rect(f["xy"], r, gx, gy)
    # keys in Gaussian order, tiles in row-major rect order
    keys, vals = [], []
    for i in vis:
        for ty in range(y0[i], y1[i]):
            for tx in range(x0[i], x1[i]):
                keys.append((ty * gx + tx) << 32 | int(f["gdepth"][i:i + 1].view(np.uint32)[0]))
                vals.append(i)
    keys = np.array(keys, np.uint64)
    vals = np.array(vals, np.uint32)
    np.testing.assert_array_equal(f["keys_dup"], keys)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(f["keys"], keys[order])
    np.testing.assert_array_equal(f["vals"], vals[order])
    # per-pixel brute force: visible Gaussians whose rect holds the pixel's tile, by (depth, index)
    img = np.zeros((3, H, W))
    for py in range(H):
        for px in range(W):
            t = (py // 16) * gx + px // 16
            ty, tx = divmod(t, gx)
            cand = [i for i in vis if x0[i] <= tx < x1[i] and y0[i] <= ty < y1[i]]
            cand.sort(key=lambda i: (f["gdepth"][i], i))
            T, C = 1.0, np.zeros(3)
            for i in cand:
                dx, dy = f["xy"][i, 0] - px, f["xy"][i, 1] - py
                a, b, c, o = f["conic_op"][i].astype(np.float64)
                pw = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
                if pw > 0:
                    continue
                al = min(0.99, o * math.exp(pw))
                if al < 1 / 255:
                    continue
                if T * (1 - al) < 1e-4:
                    break
                C += f["rgb"][i] * al * T
                T *= 1 - al
            img[:, py, px] = C + T * cam["bg"]
    np.testing.assert_allclose(f["color"], img, atol=2e-5)


# ---------------------------------------------------------------- kNN
def test_knn_brute_force_matches_kdtree():
    rng = np.random.default_rng(4)
    pts = rng.normal(0, 1, (3000, 3)).astype(np.float32)
    pts[10:20] = pts[0]
    d = O.knn3_brute(pts)
    dist, _ = cKDTree(pts.astype(np.float64)).query(pts.astype(np.float64), k=4)
    ref = (dist[:, 1:] ** 2).mean(1)
    np.testing.assert_allclose(d, ref, rtol=2e-6, atol=1e-12)
