"""Host-side pose-solve pieces: the SE(3) map in torch (the plain solver's
pose parametrisation and the reference the SE(3) kernels are checked
against), the error metrics, the frame lift.

Pins: the matrix exponential of the se(3) twist (scipy.linalg.expm) fixes
Exp; composition against 4x4 products; derivatives against central finite
differences in float64; the error metric against rotations of known angle.
"""
import math

import numpy as np
import pytest
import torch
from scipy.linalg import expm

from gs_hip.pose import lift_points, pose_errors, pose_matrix, se3_exp


def _twist(xi):
    x = np.asarray(xi, dtype=np.float64)
    A = np.zeros((4, 4))
    A[:3, :3] = [[0, -x[5], x[4]], [x[5], 0, -x[3]], [-x[4], x[3], 0]]
    A[:3, 3] = x[:3]
    return A


@pytest.mark.parametrize("scale", [0.0, 1e-5, 1e-3, 0.05, 0.7, 2.5])
def test_se3_exp_is_matrix_exponential(scale):
    g = torch.Generator().manual_seed(3)
    for _ in range(5):
        xi = torch.randn(6, generator=g, dtype=torch.float64) * scale
        T = pose_matrix(se3_exp(xi))
        np.testing.assert_allclose(T, expm(_twist(xi)), atol=1e-13)


def test_se3_compose_left():
    g = torch.Generator().manual_seed(4)
    xi = torch.randn(6, generator=g, dtype=torch.float64) * 0.4
    x0 = se3_exp(torch.randn(6, generator=g, dtype=torch.float64) * 0.8)
    got = pose_matrix(se3_exp(xi, x0))
    np.testing.assert_allclose(got, expm(_twist(xi)) @ pose_matrix(x0), atol=1e-13)


@pytest.mark.parametrize("scale", [0.0, 1e-4, 0.3])
def test_se3_exp_derivative_matches_fd(scale):
    g = torch.Generator().manual_seed(5)
    xi = (torch.randn(6, generator=g, dtype=torch.float64) * scale).requires_grad_(True)
    x0 = se3_exp(torch.randn(6, generator=g, dtype=torch.float64) * 0.5)
    w = torch.randn(7, generator=g, dtype=torch.float64)
    (se3_exp(xi, x0) * w).sum().backward()
    h = 1e-6
    fd = torch.zeros(6, dtype=torch.float64)
    for k in range(6):
        e = torch.zeros(6, dtype=torch.float64)
        e[k] = h
        with torch.no_grad():
            fd[k] = ((se3_exp(xi + e, x0) * w).sum() - (se3_exp(xi - e, x0) * w).sum()) / (2 * h)
    assert torch.isfinite(xi.grad).all()
    torch.testing.assert_close(xi.grad, fd, rtol=1e-7, atol=1e-8)


def test_se3_identity_gradient_closed_form():
    # at xi = 0, X0 = I: d t / d tau = I, d q_xyz / d phi = I / 2, the rest 0
    xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    w = torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0], dtype=torch.float64)
    (se3_exp(xi) * w).sum().backward()
    torch.testing.assert_close(xi.grad, torch.tensor([1.0, 2.0, 3.0, 2.0, 2.5, 3.0], dtype=torch.float64))


def test_pose_errors_known_angle():
    ang = 0.123
    q = [0.0, math.sin(ang / 2), 0.0, math.cos(ang / 2)]
    a = torch.tensor([0.0, 0.0, 0.0] + q, dtype=torch.float64)
    b = torch.tensor([0.3, -0.4, 0.0, 0.0, 0.0, 0.0, 1.0], dtype=torch.float64)
    rot, trans = pose_errors(a, torch.tensor([0.0, 0, 0, 0, 0, 0, 1], dtype=torch.float64))
    assert abs(rot - ang) < 1e-12 and trans < 1e-15
    rot, trans = pose_errors(b, torch.tensor([0.0, 0, 0, 0, 0, 0, 1], dtype=torch.float64))
    assert rot < 1e-7 and abs(trans - 0.5) < 1e-12


def test_lift_points_pinhole_and_supergrid():
    xyz, rgb, cam, sup = lift_points(1, 64, 36, None)
    assert xyz.shape == (64 * 36, 3) and not sup
    fx = 64 / (2 * math.tan(math.radians(30)))
    # pixel (u, v) = (10, 7) lies on the ray through it
    p = xyz[7 * 64 + 10]
    assert abs(float(p[0] / p[2]) - (10 - 32) / fx) < 1e-6
    assert abs(float(p[1] / p[2]) - (7 - 18) / fx) < 1e-6
    assert float(xyz[:, 2].min()) > 1.0
    xyz2, _, _, sup2 = lift_points(1, 64, 36, 3000)
    assert sup2 and xyz2.shape == (3000, 3)
    xyz3, _, _, _ = lift_points(1, 64, 36, 1000)
    assert xyz3.shape == (1000, 3)
    # same seed, same points
    xyz4, _, _, _ = lift_points(1, 64, 36, 1000)
    assert torch.equal(xyz3, xyz4)


def test_serving_request_stream_and_percentiles():
    from gs_hip.serve import Request, make_requests, summarize
    a = make_requests(200, rate=10.0, n_scenes=16, seed=5)
    b = make_requests(200, rate=10.0, n_scenes=16, seed=5)
    assert [(r.kind, r.arrival, r.scene, r.W) for r in a] == [(r.kind, r.arrival, r.scene, r.W) for r in b]
    arr = [r.arrival for r in a]
    assert arr == sorted(arr)
    frac = sum(r.kind == "render" for r in a) / len(a)
    assert 0.6 < frac < 0.8
    # mean inter-arrival ~ 1 / rate
    assert abs(arr[-1] / len(a) - 0.1) < 0.03
    done = []
    for i in range(100):
        r = Request(rid=i, kind="render", arrival=0.0)
        r.done = (i + 1) / 100.0
        done.append(r)
    s = summarize(done, 2.0)
    assert s["p50_s"] == 0.5 and s["p99_s"] == 0.99 and s["throughput_rps"] == 50.0
