"""The lietorch-compatible package (cf-3dgs_amd/lietorch) on the CPU.

Pins:
  * SE3.exp against scipy's matrix exponential of the 4x4 twist (both the
    Taylor and the closed-form branch), and SO3.exp against scipy's rotation
    vector;
  * log as the inverse of exp, inv as the group inverse, composition and act
    against 4x4 matrix products;
  * LieGroupParameter: zero tangent, retr() = Exp(xi) * group, gradients
    equal to central finite differences in float64, stepped by
    torch.optim.Adam as a plain parameter;
  * the same pose as ``gs_hip.pose.se3_exp``, the map of the renderer's
    fused-pose path;
  * the reference's call shapes (scene/gaussian_model_cf.py:123-147,328-340,
    368-373) run unchanged.
"""
import numpy as np
import pytest
import torch
from scipy.linalg import expm
from scipy.spatial.transform import Rotation

from gs_hip.pose import se3_exp
from lietorch import SE3, SO3, LieGroupParameter, Sim3


def _twist(xi):
    x = np.asarray(xi, dtype=np.float64)
    A = np.zeros((4, 4))
    A[:3, :3] = [[0, -x[5], x[4]], [x[5], 0, -x[3]], [-x[4], x[3], 0]]
    A[:3, 3] = x[:3]
    return A


@pytest.mark.parametrize("scale", [0.0, 1e-5, 1e-3, 0.3, 2.0])
def test_se3_exp_is_the_matrix_exponential(scale):
    g = torch.Generator().manual_seed(1)
    xi = torch.randn(8, 6, generator=g, dtype=torch.float64) * scale
    M = SE3.exp(xi).matrix().numpy()
    for k in range(8):
        np.testing.assert_allclose(M[k], expm(_twist(xi[k].numpy())), atol=1e-13)


def test_so3_exp_is_the_rotation_vector():
    g = torch.Generator().manual_seed(2)
    phi = torch.randn(16, 3, generator=g, dtype=torch.float64)
    q = SO3.exp(phi).data.numpy()
    ref = Rotation.from_rotvec(phi.numpy()).as_quat()  # (x, y, z, w), w >= 0
    np.testing.assert_allclose(q * np.sign(q[:, 3:]), ref * np.sign(ref[:, 3:]), atol=1e-14)
    np.testing.assert_allclose(SO3.exp(phi).matrix().numpy()[:, :3, :3], Rotation.from_rotvec(phi.numpy()).as_matrix(),
                               atol=1e-14)
    small = phi * (2.5 / phi.norm(dim=1, keepdim=True).clamp_min(2.5))  # angles below pi: log inverts exp
    np.testing.assert_allclose(SO3.exp(small).log().numpy(), small.numpy(), atol=1e-12)


def test_log_inv_compose_act_against_matrices():
    g = torch.Generator().manual_seed(3)
    xa = torch.randn(5, 6, generator=g, dtype=torch.float64) * 0.8
    xb = torch.randn(5, 6, generator=g, dtype=torch.float64) * 0.8
    A, B = SE3.exp(xa), SE3.exp(xb)
    np.testing.assert_allclose(A.log().numpy(), xa.numpy(), atol=1e-12)
    np.testing.assert_allclose((A * B).matrix().numpy(), (A.matrix() @ B.matrix()).numpy(), atol=1e-13)
    np.testing.assert_allclose((A * A.inv()).matrix().numpy(), np.broadcast_to(np.eye(4), (5, 4, 4)), atol=1e-13)
    p = torch.randn(5, 3, generator=g, dtype=torch.float64)
    ph = torch.cat([p, torch.ones(5, 1, dtype=torch.float64)], 1)
    want = (A.matrix() @ ph[..., None])[..., 0]
    np.testing.assert_allclose(A.act(p).numpy(), want[:, :3].numpy(), atol=1e-13)
    np.testing.assert_allclose(A.act(ph).numpy(), want.numpy(), atol=1e-13)
    np.testing.assert_allclose((A * p).numpy(), want[:, :3].numpy(), atol=1e-13)
    # a [1] group acts on [P, 3] points by broadcasting, as get_xyz uses it
    X = SE3.exp(xa[:1])
    pts = torch.randn(100, 3, generator=g, dtype=torch.float64)
    np.testing.assert_allclose(X.act(pts).numpy(), (pts @ X.matrix()[0, :3, :3].T + X.matrix()[0, :3, 3]).numpy(),
                               atol=1e-13)


def test_same_pose_as_the_fused_path_map():
    g = torch.Generator().manual_seed(4)
    xi = torch.randn(6, 6, generator=g, dtype=torch.float64) * 0.5
    X0 = SE3.exp(torch.randn(6, 6, generator=g, dtype=torch.float64))
    np.testing.assert_allclose(SE3(X0.data).retr(xi).data.numpy(), se3_exp(xi, X0.data).numpy(), atol=1e-14)


def test_lie_group_parameter_gradient_is_the_derivative():
    g = torch.Generator().manual_seed(5)
    X0 = SE3.exp(torch.randn(1, 6, generator=g, dtype=torch.float64) * 0.3)
    P = LieGroupParameter(X0)
    assert P.is_leaf and P.requires_grad and tuple(P.shape) == (1, 6) and float(P.detach().abs().sum()) == 0.0
    np.testing.assert_allclose(P.retr().data.detach().numpy(), X0.data.numpy(), atol=1e-15)
    pts = torch.randn(200, 3, generator=g, dtype=torch.float64)
    w = torch.randn(200, 3, generator=g, dtype=torch.float64)

    def f(xi):
        return (torch.sin(SE3.exp(xi).__mul__(X0).act(pts)) * w).sum()

    for xi0 in (torch.zeros(1, 6, dtype=torch.float64), torch.randn(1, 6, generator=g, dtype=torch.float64) * 0.4):
        with torch.no_grad():
            P.copy_(xi0)
        P.grad = None
        (torch.sin(P.retr().act(pts)) * w).sum().backward()
        h = 1e-6
        fd = torch.zeros(6, dtype=torch.float64)
        for k in range(6):
            e = torch.zeros(1, 6, dtype=torch.float64)
            e[0, k] = h
            fd[k] = (f(xi0 + e) - f(xi0 - e)) / (2 * h)
        np.testing.assert_allclose(P.grad[0].numpy(), fd.numpy(), rtol=1e-6, atol=1e-7)


def test_reference_call_shapes_and_adam():
    """init_RT(None), get_xyz, get_RT and training_setup_fix_position's Adam
    (scene/gaussian_model_cf.py:123-147,303-340), verbatim in shape."""
    pose_init = torch.as_tensor([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]).requires_grad_(True)
    Ps = [LieGroupParameter(SE3(pose_init[None]))]
    xyz = torch.randn(50, 3)
    out = Ps[0].retr().act(xyz.clone())
    torch.testing.assert_close(out, xyz, rtol=0, atol=1e-7)
    assert Ps[0].retr().matrix().squeeze().shape == (4, 4)
    opt = torch.optim.Adam([{"params": [Ps[0]], "lr": 1e-3, "name": "R"}], lr=0.0, eps=1e-15)
    target = SE3.exp(torch.tensor([[0.05, -0.02, 0.03, 0.02, 0.01, -0.03]])).act(xyz)
    first = None
    for it in range(200):
        loss = ((Ps[0].retr().act(xyz.clone()) - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        opt.step()
        first = float(loss) if first is None else first
        assert type(Ps[0]) is LieGroupParameter
    assert float(loss) < 0.05 * first
    # update_RT_seq: a new parameter with its group reassigned (:368-373)
    pose = torch.tensor([0.1, 0.2, 0.3, 0.0, 0.0, 0.0, 1.0])
    Ps[0] = LieGroupParameter(SE3(pose[None]))
    Ps[0].group = SE3(pose[None])
    torch.testing.assert_close(Ps[0].retr().act(torch.zeros(1, 3)), pose[None, :3])


def test_sim3_is_not_provided():
    with pytest.raises(NotImplementedError):
        Sim3(torch.zeros(1, 8))
    with pytest.raises(ValueError):
        SE3(torch.zeros(1, 6))
