"""Pins the oracle's less-travelled branches, each with a check a plausible
slip would fail:

* the 1.3 tan(fov) clamp of the EWA Jacobian (a Gaussian beyond it whose
  rect still reaches the image): the 2-D covariance against an independent
  numpy EWA with the clamped Jacobian (and a check that the unclamped one
  would differ), and fp64 central differences along the clamped axis, where
  the reference masks the Jacobian's term;
* ``scale_modifier`` != 1: the closed-form radius of an on-axis isotropic
  Gaussian, and the reference's convention that dL/dscale is the derivative
  with respect to the modified scale (finite differences of the raw scale
  equal ``scale_modifier`` times it), while the rotation and mean gradients
  stay plain derivatives;
* the SSIM of the loss (trainer/losses.py:143-209): window taps sum to 1 with
  a spread of about 1.5 px, SSIM(x, x) = 1, the per-pixel closed form of two
  constant images under zero padding, and an independent scipy evaluation.
"""
import math

import numpy as np
import pytest
import torch
from scipy import ndimage

from oracle import oracle as O
from test_oracle_closed_forms import cam_dict, one
from test_oracle_gradients import fd_check


def _ewa_conic(means, scales, cam, clamp=True):
    """Independent EWA: view = identity, Sigma = diag(s^2) (rotation identity)."""
    W, H = cam["W"], cam["H"]
    tx, ty = cam["tanfovx"], cam["tanfovy"]
    fx, fy = W / (2 * tx), H / (2 * ty)
    x, y, z = means.T
    if clamp:
        x = np.clip(x / z, -1.3 * tx, 1.3 * tx) * z
        y = np.clip(y / z, -1.3 * ty, 1.3 * ty) * z
    J = np.zeros((len(z), 2, 3))
    J[:, 0, 0] = fx / z
    J[:, 0, 2] = -fx * x / z ** 2
    J[:, 1, 1] = fy / z
    J[:, 1, 2] = -fy * y / z ** 2
    Sig = np.einsum("pi,ij->pij", scales ** 2, np.eye(3))
    C = J @ Sig @ J.transpose(0, 2, 1) + 0.3 * np.eye(2)
    return np.linalg.inv(C), C


def test_frustum_clamp_covariance_matches_independent_ewa():
    cam = cam_dict(W=64, H=48)
    tx, ty = cam["tanfovx"], cam["tanfovy"]
    z = 4.0
    # beyond 1.3 tan(fov) to the right, to the upper left and below, yet large enough to reach the image
    means = np.array([[1.42 * tx * z, 0.1, z], [-1.36 * tx * z, -1.4 * ty * z, z], [0.2, 1.5 * ty * z, z],
                      [0.3, -0.2, z]])
    scales = np.array([[0.5, 0.4, 0.6], [0.6, 0.5, 0.7], [0.5, 0.6, 0.4], [0.2, 0.3, 0.25]])
    sc = one(means, scales)
    g = O.preprocess(sc, cam, np.float64)
    assert (g["radii"] > 0).all() and (g["tiles"] > 0).all(), "every test Gaussian must reach the image"
    Ci, C = _ewa_conic(means, scales, cam, clamp=True)
    np.testing.assert_allclose(g["conic_op"][:, 0], Ci[:, 0, 0], rtol=1e-9)
    np.testing.assert_allclose(g["conic_op"][:, 1], Ci[:, 0, 1], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(g["conic_op"][:, 2], Ci[:, 1, 1], rtol=1e-9)
    mid = 0.5 * (C[:, 0, 0] + C[:, 1, 1])
    r = np.ceil(3 * np.sqrt(mid + np.sqrt(np.maximum(0.1, mid * mid - np.linalg.det(C)))))
    np.testing.assert_array_equal(g["radii"], r.astype(np.int32))
    # the clamp matters for the first three (an unclamped Jacobian gives another covariance)
    Cu, _ = _ewa_conic(means, scales, cam, clamp=False)
    rel = np.abs(Cu[:, 0, 0] - Ci[:, 0, 0]) / Ci[:, 0, 0] + np.abs(Cu[:, 1, 1] - Ci[:, 1, 1]) / Ci[:, 1, 1]
    assert (rel[:3] > 0.02).all() and rel[3] < 1e-12


def _clamp_scene():
    """A small in-frustum scene plus Gaussians beyond the horizontal clamp
    whose splats cover the image's right edge."""
    rng = np.random.default_rng(7)
    cam = cam_dict(W=48, H=40, bg=(0.1, 0.2, 0.3), deg=0)
    tx = cam["tanfovx"]
    P_in, P_out = 12, 4
    z_out = rng.uniform(3.6, 4.4, P_out)
    means = np.concatenate([
        np.stack([rng.uniform(-0.8, 0.8, P_in), rng.uniform(-0.6, 0.6, P_in), rng.uniform(3.5, 4.5, P_in)], 1),
        np.stack([rng.uniform(1.34, 1.45, P_out) * tx * z_out, rng.uniform(-0.4, 0.4, P_out), z_out], 1)])
    scales = np.concatenate([rng.uniform(0.12, 0.3, (P_in, 3)), rng.uniform(0.45, 0.6, (P_out, 3))])
    q = rng.normal(0, 1, (P_in + P_out, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    sc = dict(means=means, scales=scales, rotations=q, opacities=rng.uniform(0.3, 0.85, P_in + P_out),
              colors=rng.uniform(0.2, 0.9, (P_in + P_out, 3)))
    return sc, cam, P_in


def _loss(scene, cam, w):
    f = O.forward(scene, cam, np.float64)
    return float((w[0] * f["color"]).sum() + (w[1] * f["depth"]).sum() + (w[2] * f["alpha"]).sum())


def _fd(scene, cam, key, k, w, h=1e-6):
    """Central difference of entry k of scene[key] at steps h and h/2 (None
    when they disagree: a threshold decision flipped inside the stencil)."""
    base = np.array(scene[key], np.float64)
    vals = []
    for hh in (h, h / 2):
        p = base.reshape(-1).copy()
        p[k] += hh
        scene[key] = p.reshape(base.shape)
        lp = _loss(scene, cam, w)
        p[k] -= 2 * hh
        scene[key] = p.reshape(base.shape)
        lm = _loss(scene, cam, w)
        vals.append((lp - lm) / (2 * hh))
    scene[key] = base
    if abs(vals[0] - vals[1]) > 1e-4 * max(abs(vals[1]), 1e-8) + 1e-9:
        return None
    return vals[1]


def test_frustum_clamp_gradient_is_masked_along_the_clamped_axis():
    """Beyond the clamp the Jacobian no longer depends on the view-space x
    (it is evaluated at 1.3 tan(fovx) z), so its term of dL/dx must vanish;
    the reference masks it.  Central differences in fp64 of the x coordinate
    of every clamped Gaussian, and of every coordinate of the in-frustum
    ones, must equal the analytic gradient."""
    sc, cam, P_in = _clamp_scene()
    f = O.forward(sc, cam, np.float64)
    P = sc["means"].shape[0]
    assert (f["radii"][P_in:] > 0).all()
    rng = np.random.default_rng(1)
    w = (rng.normal(0, 1, (3, cam["H"], cam["W"])), rng.normal(0, 1, (1, cam["H"], cam["W"])),
         rng.normal(0, 1, (1, cam["H"], cam["W"])))
    g = O.backward(sc, cam, f, w[0], w[1], w[2], np.float64)
    ga = g["means3D"].reshape(-1)
    scale = np.abs(ga).max()
    # the clamped Gaussians carry real weight in the loss
    assert np.abs(g["means3D"][P_in:, 0]).min() > 1e-3 * scale
    entries = [3 * i for i in range(P_in, P)] + list(range(3 * P_in))
    used = 0
    for k in entries:
        fd = _fd(sc, cam, "means", k, w)
        if fd is None:
            continue
        used += 1
        assert abs(ga[k] - fd) <= 1e-4 * max(abs(fd), 1e-3 * scale), (k, ga[k], fd)
    assert used >= 0.9 * len(entries)


def test_scale_modifier_radius_closed_form():
    cam = cam_dict()
    cam["scale_modifier"] = 0.5
    W = cam["W"]
    z, s = 4.0, 0.4
    g = O.preprocess(one([0.0, 0.0, z], [s, s, s]), cam, np.float64)
    f = W / (2 * cam["tanfovx"])
    lam = f * f * (0.5 * s) ** 2 / (z * z) + 0.3
    assert g["radii"][0] == math.ceil(3 * math.sqrt(lam + math.sqrt(0.1)))
    np.testing.assert_allclose(g["conic_op"][0, :3], [1 / lam, 0.0, 1 / lam], rtol=1e-9, atol=1e-12)
    # and a scale of 0.5 s with modifier 1 is the same splat
    cam1 = cam_dict()
    g1 = O.preprocess(one([0.0, 0.0, z], [0.5 * s] * 3), cam1, np.float64)
    np.testing.assert_allclose(g1["conic_op"], g["conic_op"], rtol=1e-12)
    assert g1["radii"][0] == g["radii"][0]


def test_scale_modifier_gradient_convention():
    """dL/dscale is reported with respect to the modified scale
    (scale_modifier * scale): finite differences of the raw scale equal
    scale_modifier times the reported gradient.  Rotations and means get
    their plain derivatives."""
    from test_oracle_gradients import loss_and_weights, small_scene
    scene, cam = small_scene(11)
    mod = 0.5
    cam["scale_modifier"] = mod
    scene["scales"] = scene["scales"] * 1.8  # the same splat sizes as the unmodified scene, roughly
    w = loss_and_weights(11, cam["H"], cam["W"])
    f = O.forward(scene, cam, np.float64)
    g = O.backward(scene, cam, f, w[0], w[1], w[2], np.float64)
    fd_check(scene, cam, "scales", mod * g["scales"], w, seed=11)
    fd_check(scene, cam, "rotations", g["rotations"], w, seed=11)
    fd_check(scene, cam, "means", g["means3D"], w, seed=11)
    # the reported gradient is not the raw-scale derivative (the convention is visible)
    with pytest.raises(AssertionError):
        fd_check(scene, cam, "scales", g["scales"], w, seed=11)


# ---------------------------------------------------------------- SSIM
def _taps():
    from gs_hip.loss import gaussian_window
    return gaussian_window(11, 1.5, 1)[0, 0].double().numpy()


def test_ssim_window_sums_to_one_with_sigma_1_5():
    w2 = _taps()
    assert w2.shape == (11, 11)
    assert abs(w2.sum() - 1.0) < 1e-6
    k = np.arange(-5, 6)
    wx, wy = w2.sum(0), w2.sum(1)
    np.testing.assert_allclose(w2, np.outer(wy, wx), rtol=1e-6)  # separable
    assert abs((wx * k).sum()) < 1e-7 and abs((wy * k).sum()) < 1e-7  # centred
    var = (wx * k * k).sum()
    # a Gaussian of sigma 1.5 truncated at +-5 taps: variance just under 1.5^2
    assert 2.20 < var < 2.25, var
    assert np.all(np.diff(wx[:6]) > 0)


def test_ssim_of_an_image_with_itself_is_one():
    from gs_hip.loss import ssim
    g = torch.Generator().manual_seed(4)
    x = torch.rand((3, 33, 47), generator=g, dtype=torch.float64)
    assert float(ssim(x, x)) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("a,b", [(0.3, 0.7), (0.9, 0.1), (0.5, 0.5)])
def test_ssim_constant_images_closed_form(a, b):
    """Two constant images under zero padding: with s the window mass inside
    the image at a pixel, mu = a s, sigma^2 = a^2 s (1 - s), sigma_12 =
    a b s (1 - s); SSIM follows per pixel."""
    from gs_hip.loss import ssim
    H, W = 17, 23
    x = torch.full((3, H, W), a, dtype=torch.float64)
    y = torch.full((3, H, W), b, dtype=torch.float64)
    w2 = _taps()
    s = np.zeros((H, W))  # window mass that falls inside the image at each pixel
    for i in range(H):
        for j in range(W):
            s[i, j] = w2[max(0, 5 - i):min(11, H + 5 - i), max(0, 5 - j):min(11, W + 5 - j)].sum()
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    v = s * (1 - s)
    m = ((2 * a * b * s * s + C1) * (2 * a * b * v + C2)) / (((a * a + b * b) * s * s + C1) * ((a * a + b * b) * v + C2))
    assert float(ssim(x, y)) == pytest.approx(float(m.mean()), rel=1e-9)
    # interior pixels (full window inside) reduce to (2ab + C1) / (a^2 + b^2 + C1)
    assert s[8, 11] == pytest.approx(1.0, abs=1e-6)  # (float32 taps)


def test_ssim_matches_scipy_correlation():
    """Independent evaluation with scipy.ndimage (zero-padded correlation)."""
    from gs_hip.loss import ssim
    g = torch.Generator().manual_seed(5)
    x = torch.rand((3, 26, 31), generator=g, dtype=torch.float64)
    y = (x + 0.2 * torch.rand((3, 26, 31), generator=g, dtype=torch.float64)).clamp(0, 1)
    w2 = _taps()
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    tot = []
    for c in range(3):
        X, Y = x[c].numpy(), y[c].numpy()
        f = lambda im: ndimage.correlate(im, w2, mode="constant", cval=0.0)  # noqa: E731
        m1, m2 = f(X), f(Y)
        v1, v2, v12 = f(X * X) - m1 * m1, f(Y * Y) - m2 * m2, f(X * Y) - m1 * m2
        tot.append(((2 * m1 * m2 + C1) * (2 * v12 + C2)) / ((m1 * m1 + m2 * m2 + C1) * (v1 + v2 + C2)))
    assert float(ssim(x, y)) == pytest.approx(float(np.mean(tot)), rel=1e-12)


# ---------------------------------------------------------------- the CF pipeline's camera
def _co3d_cam(W=56, H=44, deg=0, R=None, t=None):
    from gs_hip.camera import make_camera_co3d, oracle_camera
    K = np.array([[61.0, 0.0, 0.37 * W], [0.0, 58.0, 0.61 * H], [0.0, 0.0, 1.0]])
    cam = make_camera_co3d(np.eye(3) if R is None else R, np.zeros(3) if t is None else t, K, W, H)
    return oracle_camera(cam, np.array([0.1, 0.2, 0.3], np.float32), deg), K, cam


def test_co3d_world_to_view_and_principal_point():
    """getWorld2View3 is [R t; 0 1] (R world-to-camera); a point at camera
    coordinates (X, Y, Z) lands on the pixel centre (fx X/Z + cx - 0.5,
    fy Y/Z + cy - 0.5) with an off-centre principal point."""
    from gs_hip.camera import world_to_view3
    from scipy.spatial.transform import Rotation
    R = Rotation.from_rotvec([0.1, -0.2, 0.05]).as_matrix()
    t = np.array([0.3, -0.1, 0.5])
    V = world_to_view3(R, t)
    np.testing.assert_allclose(V[:3, :3], R, atol=1e-7)
    np.testing.assert_allclose(V[:3, 3], t, atol=1e-7)
    camd, K, cam = _co3d_cam(R=R, t=t)
    np.testing.assert_allclose(cam.camera_center.numpy(), -R.T @ t, atol=1e-5)
    rng = np.random.default_rng(0)
    Xc = np.stack([rng.uniform(-0.8, 0.8, 30), rng.uniform(-0.6, 0.6, 30), rng.uniform(2.5, 5.0, 30)], 1)
    means = (Xc - t) @ R  # world points: R p + t = Xc
    g = O.preprocess(one(means, [0.05, 0.05, 0.05]), camd, np.float64)
    assert (g["radii"] > 0).all()
    np.testing.assert_allclose(g["xy"][:, 0], K[0, 0] * Xc[:, 0] / Xc[:, 2] + K[0, 2] - 0.5, atol=2e-4)
    np.testing.assert_allclose(g["xy"][:, 1], K[1, 1] * Xc[:, 1] / Xc[:, 2] + K[1, 2] - 0.5, atol=2e-4)
    np.testing.assert_allclose(g["gdepth"], Xc[:, 2], rtol=1e-6)
    # the rasteriser's EWA focal is fx: W / (2 tan(FoVx / 2)) with FoVx = focal2fov(fx, W)
    assert camd["W"] / (2 * camd["tanfovx"]) == pytest.approx(K[0, 0], rel=1e-12)


def test_fd_means_with_off_centre_principal_point():
    """The projection gradient with the principal-point column of the
    is_co3d projection (zero in the legacy camera) against fp64 central
    differences."""
    from test_oracle_gradients import small_scene
    from scipy.spatial.transform import Rotation
    scene, _ = small_scene(12, W=56, H=44)
    camd, _, _ = _co3d_cam(deg=3, R=Rotation.from_rotvec([0.02, -0.03, 0.01]).as_matrix(),
                           t=np.array([0.05, -0.02, 0.1]))
    assert abs(camd["proj"].reshape(4, 4)[2, 0]) > 0.1  # the principal-point term is live
    rng = np.random.default_rng(12)
    w = (rng.normal(0, 1, (3, 44, 56)), rng.normal(0, 1, (1, 44, 56)), rng.normal(0, 1, (1, 44, 56)))
    f = O.forward(scene, camd, np.float64)
    assert f["num_rendered"] > 0
    g = O.backward(scene, camd, f, w[0], w[1], w[2], np.float64)
    fd_check(scene, camd, "means", g["means3D"], w, seed=12)
    fd_check(scene, camd, "shs", g["shs"], w, nsamp=40, seed=12)
