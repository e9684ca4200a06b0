"""Pins the oracle's analytic backward with fp64 central finite differences.

Loss L = sum(wc * colour) + sum(wd * depth) + sum(wa * alpha) with random
weights; every parameter group of the rasteriser (means, scales, rotations,
opacities, SH, precomputed colours and covariances, and the rigid pose)
is differentiated numerically on a small scene.  The scenes avoid the places
where the reference's gradient is knowingly not the derivative (the 0.99
alpha clamp, the 1.3 tan(fov) clamp): opacities <= 0.9 and all means well
inside the frustum.  An entry is compared only where two step sizes (h, h/2)
agree -- a tile-rect or threshold decision flipping inside the stencil makes
the difference quotient meaningless there -- and at least 90 % of the sampled
entries of every group must be comparable.
"""
import math

import numpy as np
import pytest

from gs_hip import scenes
from gs_hip.camera import make_camera, oracle_camera
from oracle import oracle as O


def small_scene(seed, P=40, W=48, H=40, deg=3):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-0.8, 0.8, P), rng.uniform(-0.6, 0.6, P), rng.uniform(3.5, 4.5, P)], 1)
    scales = rng.uniform(0.12, 0.3, (P, 3))
    q = rng.normal(0, 1, (P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opac = rng.uniform(0.2, 0.9, P)
    shs = rng.normal(0, 0.15, (P, 16, 3))
    shs[:, 0, :] += 1.0  # keep colours away from the max(0, .) clamp
    fov = math.radians(60)
    cam = make_camera(np.eye(3), np.zeros(3), fov, fov * H / W, W, H)
    camd = oracle_camera(cam, np.array([0.1, 0.2, 0.3], np.float32), deg)
    return dict(means=means, scales=scales, rotations=q, opacities=opac, shs=shs), camd


def loss_and_weights(seed, H, W):
    rng = np.random.default_rng(1000 + seed)
    return rng.normal(0, 1, (3, H, W)), rng.normal(0, 1, (1, H, W)), rng.normal(0, 1, (1, H, W))


def L(scene, cam, w):
    f = O.forward(scene, cam, np.float64)
    return float((w[0] * f["color"]).sum() + (w[1] * f["depth"]).sum() + (w[2] * f["alpha"]).sum())


def fd_check(scene, cam, key, analytic, w, nsamp=30, h=1e-6, seed=0, cam_key=False):
    src = cam if cam_key else scene
    base = np.array(src[key], np.float64)
    flat = base.reshape(-1)
    rng = np.random.default_rng(seed)
    idx = rng.choice(flat.size, size=min(nsamp, flat.size), replace=False)
    ok, used = 0, 0
    worst = 0.0
    for k in idx:
        vals = []
        for hh in (h, h / 2):
            p = flat.copy()
            p[k] += hh
            src[key] = p.reshape(base.shape)
            lp = L(scene, cam, w)
            p[k] -= 2 * hh
            src[key] = p.reshape(base.shape)
            lm = L(scene, cam, w)
            vals.append((lp - lm) / (2 * hh))
        src[key] = base
        fd1, fd2 = vals
        scale = max(abs(fd1), abs(fd2), 1e-8)
        if abs(fd1 - fd2) > 1e-4 * scale + 1e-9:
            continue  # a discrete decision changed inside the stencil
        used += 1
        a = float(analytic.reshape(-1)[k])
        # relative error, floored at 1e-3 of the group's largest gradient so
        # entries near zero are judged against the group's scale
        err = abs(a - fd2) / max(abs(fd2), 1e-3 * (np.abs(analytic).max() + 1e-12))
        worst = max(worst, err)
        if err < 1e-4:
            ok += 1
    assert used >= 0.9 * len(idx), (key, used, len(idx))
    assert ok == used, (key, ok, used, worst)


@pytest.mark.parametrize("seed", [0, 1])
def test_fd_all_parameter_groups(seed):
    scene, cam = small_scene(seed)
    w = loss_and_weights(seed, cam["H"], cam["W"])
    f = O.forward(scene, cam, np.float64)
    assert f["num_rendered"] > 0
    g = O.backward(scene, cam, f, w[0], w[1], w[2], np.float64)
    fd_check(scene, cam, "means", g["means3D"], w, seed=seed)
    fd_check(scene, cam, "scales", g["scales"], w, seed=seed)
    fd_check(scene, cam, "rotations", g["rotations"], w, seed=seed)
    fd_check(scene, cam, "opacities", g["opacities"], w, seed=seed)
    fd_check(scene, cam, "shs", g["shs"], w, nsamp=60, seed=seed)


def test_fd_colors_and_cov3D_precomp():
    scene, cam = small_scene(3)
    rng = np.random.default_rng(3)
    P = scene["means"].shape[0]
    s2 = scene["scales"] ** 2
    A = rng.normal(0, 0.02, (P, 3, 3))
    S = np.einsum("pij,pkj->pik", A, A) + np.stack([np.diag(v) for v in s2])
    cov = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)
    sc2 = dict(means=scene["means"], opacities=scene["opacities"], colors=rng.uniform(0.1, 0.9, (P, 3)), cov3D=cov)
    w = loss_and_weights(3, cam["H"], cam["W"])
    f = O.forward(sc2, cam, np.float64)
    g = O.backward(sc2, cam, f, w[0], w[1], w[2], np.float64)
    fd_check(sc2, cam, "colors", g["colors"], w)
    fd_check(sc2, cam, "cov3D", g["cov3D"], w)
    fd_check(sc2, cam, "means", g["means3D"], w)


def test_fd_pose():
    scene, cam = small_scene(4)
    cam["pose7"] = np.array([0.03, -0.02, 0.05, 0.02, -0.03, 0.01, 0.99])
    w = loss_and_weights(4, cam["H"], cam["W"])
    f = O.forward(scene, cam, np.float64)
    g = O.backward(scene, cam, f, w[0], w[1], w[2], np.float64)
    fd_check(scene, cam, "pose7", g["pose7"], w, nsamp=7, cam_key=True)
    fd_check(scene, cam, "means", g["means3D"], w)


def test_pose_gradient_closed_form_at_identity():
    """At the identity pose dL/dt = sum_i g_i and, with q = (x, y, z, w=1),
    dL/d(x,y,z) = 2 sum_i p_i x g_i (p' = p + 2 v x p to first order);
    dL/dw = 0 (the normalisation removes the radial direction)."""
    scene, cam = small_scene(5)
    w = loss_and_weights(5, cam["H"], cam["W"])
    f0 = O.forward(scene, cam, np.float64)
    g0 = O.backward(scene, cam, f0, w[0], w[1], w[2], np.float64)
    cam["pose7"] = np.array([0, 0, 0, 0, 0, 0, 1.0])
    f = O.forward(scene, cam, np.float64)
    g = O.backward(scene, cam, f, w[0], w[1], w[2], np.float64)
    gi = g0["means3D"]
    np.testing.assert_allclose(g["means3D"], gi, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(g["pose7"][:3], gi.sum(0), rtol=1e-10, atol=1e-12)
    rot = 2 * np.cross(scene["means"], gi).sum(0)
    np.testing.assert_allclose(g["pose7"][3:6], rot, rtol=1e-10, atol=1e-12)
    assert abs(g["pose7"][6]) < 1e-12
