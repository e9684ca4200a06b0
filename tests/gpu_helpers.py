"""Shared helpers for the GPU parity tests (HIP path vs the CPU oracle)."""
from __future__ import annotations

import numpy as np
import torch

from gs_hip import camera as gcam
from gs_hip import raster
from oracle import oracle as O

DEV = torch.device("cuda:0")


def run_gpu(sc: dict, sh_degree: int, pose=None, colors=None, cov3D=None, with_sh=True, binning=None):
    cam = sc["cam"].to(DEV)
    rs = gcam.raster_settings(cam, sc["bg"].to(DEV), sh_degree)
    means = sc["means"].to(DEV)
    shs = sc["shs"].to(DEV) if (with_sh and colors is None) else None
    col = colors.to(DEV) if colors is not None else None
    sc_ = sc["scales"].to(DEV) if cov3D is None else None
    ro = sc["rotations"].to(DEV) if cov3D is None else None
    cv = cov3D.to(DEV) if cov3D is not None else None
    po = pose.to(DEV) if pose is not None else None
    color, radii, depth, alpha, st = raster.forward(rs, means, shs, col, sc["opacities"].to(DEV), sc_, ro, cv,
                                                    pose=po, binning=binning)
    torch.cuda.synchronize()
    P = means.shape[0]
    v = raster.views(st, P, cam.image_width, cam.image_height)
    return dict(rs=rs, color=color, radii=radii, depth=depth, alpha=alpha, st=st, views=v,
                inputs=(means, shs, col, sc["opacities"].to(DEV), sc_, ro, cv, po))


def run_oracle(sc: dict, sh_degree: int, pose=None, colors=None, cov3D=None, dtype=np.float32, with_sh=True):
    cam = gcam.oracle_camera(sc["cam"], sc["bg"], sh_degree, pose=pose)
    scene = dict(means=sc["means"].numpy(), opacities=sc["opacities"].numpy())
    if colors is not None:
        scene["colors"] = colors.numpy()
    elif with_sh:
        scene["shs"] = sc["shs"].numpy()
    if cov3D is not None:
        scene["cov3D"] = cov3D.numpy()
    else:
        scene["scales"] = sc["scales"].numpy()
        scene["rotations"] = sc["rotations"].numpy()
    return scene, cam, O.forward(scene, cam, dtype)


def np_(t):
    return t.detach().cpu().numpy()


def u64(a):
    """int64 tensor/array holding uint64 bit patterns -> numpy uint64."""
    a = np_(a) if isinstance(a, torch.Tensor) else a
    return a.view(np.uint64)


def assert_close_grad(name, got, ref, rtol=2e-3, atol_frac=2e-4):
    """Gradient tolerance: |got - ref| <= atol + rtol |ref|, atol = atol_frac * max|ref|.

    The GPU sums per-pixel contributions with wave reductions and float
    atomics in a different order from the oracle's per-thread buffers, so
    gradients agree to accumulated fp32 rounding, not bitwise."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    atol = atol_frac * scale + 1e-12
    err = np.abs(got - ref)
    bad = err > atol + rtol * np.abs(ref)
    assert not bad.any(), (f"{name}: {int(bad.sum())} of {bad.size} entries out of tolerance; "
                           f"max err {err.max():.3e}, scale {scale:.3e}")
