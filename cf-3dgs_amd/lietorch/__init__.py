"""lietorch-compatible ``SE3``, ``SO3`` and ``LieGroupParameter`` for CF-3DGS.

CF-3DGS uses lietorch in these places:

  * ``from lietorch import SO3, SE3, Sim3, LieGroupParameter``
    (scene/gaussian_model_cf.py:11);
  * ``LieGroupParameter(SE3(pose7[None]))`` (:332,337,346,354,372);
  * ``P.retr().act(xyz)`` (:127,130) and ``P.retr().matrix()`` (:140-145);
  * assigning ``P.group = SE3(...)`` (:373);
  * ``SE3(cam)`` (scene/camera_model.py:40).

lietorch itself is a C++/CUDA extension behind an unpinned git URL
(requirements.txt:14), and there is no ROCm build of it here.  This module
restates its published conventions in torch ops, on any device:

  * an SE3 element is [..., 7] = (t, q) with q = (x, y, z, w);
  * a tangent is [..., 6] = (tau, phi), translation part first;
  * Exp(tau, phi) = (V(phi) tau, quat(phi)) with
    V = I + (1 - cos th)/th^2 [phi]x + (th - sin th)/th^3 [phi]x^2;
  * composition is (t1, q1) * (t2, q2) = (t1 + R(q1) t2, q1 q2);
  * ``X.act(p)`` = R p + t, broadcast over the batch dimensions;
  * ``LieGroupParameter`` is a tensor of tangent zeros (the optimiser's
    parameter) with a base ``group``, and ``retr()`` = Exp(self) * group.

Gradients come from torch autograd through the exact map.  lietorch's
Jacobians give the same derivative, since d/dxi f(Exp(xi) X0) is unique.
The renderer's fused path (the pose passed to preprocess,
``gs_hip.raster.forward(..., pose=...)``) computes the same pose with the
``gs_se3_*`` kernels of libgs_hip.so.  It is the fast path; this module lets
the reference's code run unchanged.

``Sim3`` is imported by the reference but never constructed; here it raises.
"""
from __future__ import annotations

import torch

__all__ = ["SO3", "SE3", "Sim3", "LieGroupParameter"]


def _cross(a, b):
    return torch.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                        a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], dim=-1)


def _qmul(q1, q2):
    u1, w1 = q1[..., :3], q1[..., 3:]
    u2, w2 = q2[..., :3], q2[..., 3:]
    return torch.cat([w1 * u2 + w2 * u1 + _cross(u1, u2), w1 * w2 - (u1 * u2).sum(-1, keepdim=True)], dim=-1)


def _qrot(q, v):
    u, w = q[..., :3], q[..., 3:]
    uv = _cross(u, v)
    return v + 2.0 * w * uv + 2.0 * _cross(u, uv)


def _qconj(q):
    return torch.cat([-q[..., :3], q[..., 3:]], dim=-1)


def _qmatrix(q):
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    r = [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
         2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
         2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]
    return torch.stack(r, dim=-1).reshape(q.shape[:-1] + (3, 3))


def _so3_coeffs(phi):
    """(sin(th/2)/th, cos(th/2), (1-cos th)/th^2, (th-sin th)/th^3), with
    Taylor series below th^2 = 1e-6 so that the derivative at 0 is exact."""
    th2 = (phi * phi).sum(-1, keepdim=True)
    small = th2 < 1e-6
    th2s = torch.where(small, torch.ones_like(th2), th2)
    th = torch.sqrt(th2s)
    th4 = th2 * th2
    s = torch.where(small, 0.5 - th2 / 48.0 + th4 / 3840.0, torch.sin(0.5 * th) / th)
    w = torch.where(small, 1.0 - th2 / 8.0 + th4 / 384.0, torch.cos(0.5 * th))
    c1 = torch.where(small, 0.5 - th2 / 24.0 + th4 / 720.0, (1.0 - torch.cos(th)) / th2s)
    c2 = torch.where(small, 1.0 / 6.0 - th2 / 120.0 + th4 / 5040.0, (th - torch.sin(th)) / (th2s * th))
    return s, w, c1, c2


def _so3_log(q):
    """phi with quat(phi) = q (shortest rotation, w made non-negative)."""
    q = torch.where(q[..., 3:] < 0, -q, q)
    u, w = q[..., :3], q[..., 3:]
    n = torch.sqrt((u * u).sum(-1, keepdim=True))
    small = n < 1e-6
    ns = torch.where(small, torch.ones_like(n), n)
    # phi = 2 atan2(|u|, w) u / |u|; near 0 the series 2 u / w (1 - |u|^2 / (3 w^2))
    big = 2.0 * torch.atan2(ns, w) / ns
    ser = 2.0 / w - 2.0 * n * n / (3.0 * w * w * w)
    return torch.where(small, ser, big) * u


class _Group:
    """Shared batch behaviour of the group classes."""
    dim = 0         # storage width
    manifold_dim = 0

    def __init__(self, data: torch.Tensor):
        if isinstance(data, _Group):
            data = data.data
        if not isinstance(data, torch.Tensor) or data.shape[-1] != self.dim:
            raise ValueError(f"{type(self).__name__} expects a tensor with last dimension {self.dim}")
        self.data = data

    @property
    def shape(self):
        return self.data.shape[:-1]

    @property
    def device(self):
        return self.data.device

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def tangent_shape(self):
        return self.data.shape[:-1] + (self.manifold_dim,)

    def vec(self):
        return self.data

    def __len__(self):
        return self.data.shape[0]

    def __getitem__(self, index):
        return type(self)(self.data[index])

    def __setitem__(self, index, item):
        self.data[index] = item.data

    def to(self, *args, **kwargs):
        return type(self)(self.data.to(*args, **kwargs))

    def cpu(self):
        return type(self)(self.data.cpu())

    def cuda(self):
        return type(self)(self.data.cuda())

    def float(self):
        return type(self)(self.data.float())

    def double(self):
        return type(self)(self.data.double())

    def detach(self):
        return type(self)(self.data.detach())

    def requires_grad_(self, flag: bool = True):
        self.data.requires_grad_(flag)
        return self

    def retr(self, a):
        """Retraction Exp(a) * self (lietorch's LieGroup.retr)."""
        return type(self).exp(a) * self

    def __mul__(self, other):
        if isinstance(other, _Group):
            return self._compose(other)
        return self.act(other)

    def __repr__(self):
        return f"{type(self).__name__}({self.data!r})"


class SO3(_Group):
    """Rotations as unit quaternions [..., 4] = (x, y, z, w)."""
    dim = 4
    manifold_dim = 3

    @classmethod
    def Identity(cls, *batch, device=None, dtype=torch.float32, requires_grad=False):
        d = torch.zeros(tuple(batch) + (4,), device=device, dtype=dtype)
        d[..., 3] = 1.0
        return cls(d.requires_grad_(requires_grad))

    @classmethod
    def exp(cls, phi: torch.Tensor):
        s, w, _, _ = _so3_coeffs(phi)
        return cls(torch.cat([s * phi, w], dim=-1))

    def log(self):
        return _so3_log(self.data)

    def inv(self):
        return SO3(_qconj(self.data))

    def _compose(self, other):
        return SO3(_qmul(self.data, other.data))

    def act(self, p: torch.Tensor):
        if p.shape[-1] == 4:  # homogeneous points: rotate xyz, keep w
            return torch.cat([_qrot(self.data, p[..., :3]), p[..., 3:]], dim=-1)
        return _qrot(self.data, p)

    def matrix(self):
        R = _qmatrix(self.data)
        z3 = torch.zeros(self.data.shape[:-1] + (3, 1), device=self.device, dtype=self.dtype)
        bottom = torch.zeros(self.data.shape[:-1] + (1, 4), device=self.device, dtype=self.dtype)
        bottom[..., 0, 3] = 1.0
        return torch.cat([torch.cat([R, z3], dim=-1), bottom], dim=-2)


class SE3(_Group):
    """Rigid motions [..., 7] = (t, q) with q = (x, y, z, w)."""
    dim = 7
    manifold_dim = 6

    @classmethod
    def Identity(cls, *batch, device=None, dtype=torch.float32, requires_grad=False):
        d = torch.zeros(tuple(batch) + (7,), device=device, dtype=dtype)
        d[..., 6] = 1.0
        return cls(d.requires_grad_(requires_grad))

    @classmethod
    def InitFromVec(cls, data: torch.Tensor):
        return cls(data)

    @classmethod
    def exp(cls, xi: torch.Tensor):
        tau, phi = xi[..., :3], xi[..., 3:]
        s, w, c1, c2 = _so3_coeffs(phi)
        k1 = _cross(phi, tau)
        t = tau + c1 * k1 + c2 * _cross(phi, k1)
        return cls(torch.cat([t, s * phi, w], dim=-1))

    def log(self):
        t, q = self.data[..., :3], self.data[..., 3:]
        phi = _so3_log(q)
        _, _, c1, c2 = _so3_coeffs(phi)
        # tau = V^-1 t with V = I + c1 K + c2 K^2, K = [phi]x
        z = torch.zeros_like(phi[..., 0])
        K = torch.stack([torch.stack([z, -phi[..., 2], phi[..., 1]], -1),
                         torch.stack([phi[..., 2], z, -phi[..., 0]], -1),
                         torch.stack([-phi[..., 1], phi[..., 0], z], -1)], -2)
        eye = torch.eye(3, device=phi.device, dtype=phi.dtype).expand(K.shape)
        V = eye + c1[..., None] * K + c2[..., None] * (K @ K)
        tau = torch.linalg.solve(V, t[..., None])[..., 0]
        return torch.cat([tau, phi], dim=-1)

    def inv(self):
        t, q = self.data[..., :3], self.data[..., 3:]
        qi = _qconj(q)
        return SE3(torch.cat([-_qrot(qi, t), qi], dim=-1))

    def _compose(self, other):
        t1, q1 = self.data[..., :3], self.data[..., 3:]
        t2, q2 = other.data[..., :3], other.data[..., 3:]
        return SE3(torch.cat([t1 + _qrot(q1, t2), _qmul(q1, q2)], dim=-1))

    def translation(self):
        return self.data[..., :3]

    def rotation(self):
        return SO3(self.data[..., 3:])

    def act(self, p: torch.Tensor):
        """R p + t for [..., 3] points; [..., 4] homogeneous points get
        R p + w t (lietorch's convention for the 4-vector form)."""
        t, q = self.data[..., :3], self.data[..., 3:]
        if p.shape[-1] == 4:
            return torch.cat([_qrot(q, p[..., :3]) + p[..., 3:] * t, p[..., 3:]], dim=-1)
        return _qrot(q, p) + t

    def matrix(self):
        R = _qmatrix(self.data[..., 3:])
        t = self.data[..., :3]
        top = torch.cat([R, t[..., None]], dim=-1)
        bottom = torch.zeros(self.data.shape[:-1] + (1, 4), device=self.device, dtype=self.dtype)
        bottom[..., 0, 3] = 1.0
        return torch.cat([top, bottom], dim=-2)


class Sim3(_Group):
    """Imported by the reference (scene/gaussian_model_cf.py:11) but never
    constructed on its CF-3DGS path; not provided."""
    dim = 8
    manifold_dim = 7

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("Sim3 is not provided by this build (CF-3DGS uses SE3 only)")


class LieGroupParameter(torch.Tensor):
    """The optimiser's view of a group element: a leaf tensor of tangent
    zeros, ``group.tangent_shape``, with ``retr()`` = Exp(self) * group.
    Torch operations on it return plain tensors, so optimisers step the
    tangent like any parameter."""
    from torch._C import _disabled_torch_function_impl
    __torch_function__ = _disabled_torch_function_impl

    def __new__(cls, group: _Group, requires_grad: bool = True):
        data = torch.zeros(group.tangent_shape, device=group.device, dtype=group.dtype)
        return torch.Tensor._make_subclass(cls, data, requires_grad)

    def __init__(self, group: _Group, requires_grad: bool = True):
        self.group = group

    def retr(self):
        return self.group.retr(self)

    def log(self):
        return self.retr().log()

    def inv(self):
        return self.retr().inv()

    def act(self, p):
        return self.retr().act(p)

    def matrix(self):
        return self.retr().matrix()

    def __mul__(self, other):
        if isinstance(other, LieGroupParameter):
            return self.retr() * other.retr()
        return self.retr() * other
