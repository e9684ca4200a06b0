"""CPU oracle for the Gaussian rasteriser -- TEST INFRASTRUCTURE ONLY.

numpy/ctypes front end of ``oracle/gs_oracle.cpp`` (a C++ restatement of the
3D Gaussian Splatting rasteriser that CF-3DGS imports as
``diff_gaussian_rasterization`` at scene/gaussian_model_cf.py:28-31 and of
``simple_knn._C.distCUDA2`` used at scene/gaussian_model_cf.py:194-196).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
baseline; the product path never routes through it.

Conventions follow the reference callers:
  * ``view``/``proj`` are the 16 floats of ``world_view_transform`` /
    ``full_proj_transform`` in torch row-major order (scene/cameras.py:58-73),
    i.e. column-major world-to-camera / full projection.
  * ``shs`` is [P, M, 3] coefficient-major (scene/gaussian_model_cf.py:157-160).
  * ``pose7`` is [tx, ty, tz, qx, qy, qz, qw] (scene/gaussian_model_cf.py:330-337),
    applied to the means only (:123-133).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

BLOCK = 16


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_bin_f32.restype = ctypes.c_longlong
        _lib.orc_bin_f64.restype = ctypes.c_longlong
        _lib.orc_preprocess_f32.restype = ctypes.c_int
        _lib.orc_preprocess_f64.restype = ctypes.c_int
        _lib.orc_get_threads.restype = ctypes.c_int
    return _lib


def set_threads(n: int) -> None:
    _load().orc_set_threads(int(n))


def get_threads() -> int:
    return int(_load().orc_get_threads())


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _arr(x, dt):
    if x is None:
        return None
    return np.ascontiguousarray(np.asarray(x, dtype=dt))


def tile_grid(W: int, H: int):
    return (W + BLOCK - 1) // BLOCK, (H + BLOCK - 1) // BLOCK


def preprocess(scene: dict, cam: dict, dtype=np.float32) -> dict:
    lib = _load()
    suf = "f32" if dtype == np.float32 else "f64"
    means = _arr(scene["means"], dtype)
    P = means.shape[0]
    shs = _arr(scene.get("shs"), dtype)
    M = 0 if shs is None else shs.shape[1]
    colors = _arr(scene.get("colors"), dtype)
    opac = _arr(np.reshape(scene["opacities"], (-1,)), dtype)
    scales = _arr(scene.get("scales"), dtype)
    rots = _arr(scene.get("rotations"), dtype)
    cov_pre = _arr(scene.get("cov3D"), dtype)
    view = _arr(np.reshape(cam["view"], (16,)), dtype)
    proj = _arr(np.reshape(cam["proj"], (16,)), dtype)
    campos = _arr(np.reshape(cam["campos"], (3,)), dtype)
    pose7 = _arr(cam.get("pose7"), dtype)
    W, H = int(cam["W"]), int(cam["H"])
    out = dict(
        radii=np.zeros(P, np.int32),
        xy=np.zeros((P, 2), dtype),
        gdepth=np.zeros(P, dtype),
        conic_op=np.zeros((P, 4), dtype),
        rgb=np.zeros((P, 3), dtype),
        clamped=np.zeros((P, 3), np.uint8),
        tiles=np.zeros(P, np.uint32),
        cov3D=np.zeros((P, 6), dtype),
        means_tf=np.zeros((P, 3), dtype),
    )
    fn = getattr(lib, "orc_preprocess_" + suf)
    err = fn(
        ctypes.c_int(P), ctypes.c_int(int(cam["sh_degree"])), ctypes.c_int(M), ctypes.c_int(W), ctypes.c_int(H),
        _p(means), _p(shs), _p(colors), _p(opac), _p(scales), _p(rots), _p(cov_pre),
        ctypes.c_double(float(cam.get("scale_modifier", 1.0))), _p(view), _p(proj), _p(campos),
        ctypes.c_double(float(cam["tanfovx"])), ctypes.c_double(float(cam["tanfovy"])), _p(pose7),
        ctypes.c_int(int(bool(cam.get("prefiltered", False)))),
        _p(out["radii"]), _p(out["xy"]), _p(out["gdepth"]), _p(out["conic_op"]), _p(out["rgb"]),
        _p(out["clamped"]), _p(out["tiles"]), _p(out["cov3D"]), _p(out["means_tf"]),
    )
    out["prefilter_error"] = int(err)
    return out


def bin_and_blend(geom: dict, cam: dict, dtype=np.float32) -> dict:
    lib = _load()
    suf = "f32" if dtype == np.float32 else "f64"
    W, H = int(cam["W"]), int(cam["H"])
    gx, gy = tile_grid(W, H)
    P = geom["radii"].shape[0]
    n = int(geom["tiles"].astype(np.int64).sum())
    keys_dup = np.zeros(max(n, 1), np.uint64)
    vals_dup = np.zeros(max(n, 1), np.uint32)
    keys = np.zeros(max(n, 1), np.uint64)
    vals = np.zeros(max(n, 1), np.uint32)
    ranges = np.zeros((gx * gy, 2), np.uint32)
    got = getattr(lib, "orc_bin_" + suf)(
        ctypes.c_int(P), ctypes.c_int(W), ctypes.c_int(H), _p(geom["radii"]), _p(geom["xy"]), _p(geom["gdepth"]),
        _p(geom["tiles"]), _p(keys_dup), _p(vals_dup), _p(keys), _p(vals), _p(ranges),
    )
    assert got == n
    bg = _arr(np.reshape(cam["bg"], (3,)), dtype)
    img = dict(
        color=np.zeros((3, H, W), dtype),
        depth=np.zeros((1, H, W), dtype),
        alpha=np.zeros((1, H, W), dtype),
        final_T=np.zeros((H, W), dtype),
        n_contrib=np.zeros((H, W), np.int32),
    )
    getattr(lib, "orc_blend_" + suf)(
        ctypes.c_int(W), ctypes.c_int(H), _p(geom["xy"]), _p(geom["conic_op"]), _p(geom["rgb"]), _p(geom["gdepth"]),
        _p(bg), _p(vals), _p(ranges), _p(img["color"]), _p(img["depth"]), _p(img["alpha"]), _p(img["final_T"]),
        _p(img["n_contrib"]),
    )
    img.update(num_rendered=n, keys_dup=keys_dup[:n], vals_dup=vals_dup[:n], keys=keys[:n], vals=vals[:n],
               ranges=ranges)
    return img


def forward(scene: dict, cam: dict, dtype=np.float32) -> dict:
    """Full forward: preprocess, binning (keys, stable sort, ranges), blend.

    Per-Gaussian view depth is ``gdepth``; ``depth`` is the depth image."""
    geom = preprocess(scene, cam, dtype)
    img = bin_and_blend(geom, cam, dtype)
    out = dict(geom)
    out.update(img)
    return out


def backward(scene: dict, cam: dict, fwd: dict, dL_dcolor, dL_ddepth=None, dL_dalpha=None,
             dtype=np.float32) -> dict:
    """Analytic backward of :func:`forward` for the given output cotangents."""
    lib = _load()
    suf = "f32" if dtype == np.float32 else "f64"
    means = _arr(scene["means"], dtype)
    P = means.shape[0]
    shs = _arr(scene.get("shs"), dtype)
    M = 0 if shs is None else shs.shape[1]
    opac = _arr(np.reshape(scene["opacities"], (-1,)), dtype)
    scales = _arr(scene.get("scales"), dtype)
    rots = _arr(scene.get("rotations"), dtype)
    cov_pre = _arr(scene.get("cov3D"), dtype)
    view = _arr(np.reshape(cam["view"], (16,)), dtype)
    proj = _arr(np.reshape(cam["proj"], (16,)), dtype)
    campos = _arr(np.reshape(cam["campos"], (3,)), dtype)
    bg = _arr(np.reshape(cam["bg"], (3,)), dtype)
    pose7 = _arr(cam.get("pose7"), dtype)
    W, H = int(cam["W"]), int(cam["H"])
    g = dict(
        means2D=np.zeros((P, 3), dtype),
        colors=np.zeros((P, 3), dtype),
        opacities=np.zeros(P, dtype),
        means3D=np.zeros((P, 3), dtype),
        cov3D=np.zeros((P, 6), dtype),
        shs=np.zeros((P, max(M, 1), 3), dtype),
        scales=np.zeros((P, 3), dtype),
        rotations=np.zeros((P, 4), dtype),
        pose7=np.zeros(7, dtype),
    )
    dc = _arr(dL_dcolor, dtype)
    dd = _arr(dL_ddepth, dtype)
    da = _arr(dL_dalpha, dtype)
    vals = fwd["vals"] if fwd["num_rendered"] > 0 else np.zeros(1, np.uint32)
    getattr(lib, "orc_backward_" + suf)(
        ctypes.c_int(P), ctypes.c_int(int(cam["sh_degree"])), ctypes.c_int(M), ctypes.c_int(W), ctypes.c_int(H),
        _p(means), _p(shs), _p(opac), _p(scales), _p(rots), _p(cov_pre),
        ctypes.c_double(float(cam.get("scale_modifier", 1.0))), _p(view), _p(proj), _p(campos), _p(bg),
        ctypes.c_double(float(cam["tanfovx"])), ctypes.c_double(float(cam["tanfovy"])), _p(pose7),
        _p(fwd["radii"]), _p(_arr(fwd["xy"], dtype)), _p(_arr(fwd["gdepth"], dtype)),
        _p(_arr(fwd["conic_op"], dtype)), _p(_arr(fwd["rgb"], dtype)), _p(fwd["clamped"]),
        _p(_arr(fwd["cov3D"], dtype)), _p(_arr(fwd["means_tf"], dtype)), _p(vals), _p(fwd["ranges"]),
        _p(_arr(fwd["final_T"], dtype)), _p(fwd["n_contrib"]), _p(dc), _p(dd), _p(da),
        _p(g["means2D"]), _p(g["colors"]), _p(g["opacities"]), _p(g["means3D"]), _p(g["cov3D"]),
        _p(g["shs"] if shs is not None else None), _p(g["scales"] if scales is not None else None),
        _p(g["rotations"] if rots is not None else None), _p(g["pose7"] if pose7 is not None else None),
    )
    return g


def knn3_brute(points) -> np.ndarray:
    """Exact mean squared distance to the 3 nearest other points, O(P^2)."""
    lib = _load()
    pts = _arr(points, np.float32).reshape(-1, 3)
    out = np.zeros(pts.shape[0], np.float32)
    lib.orc_knn3_brute(ctypes.c_int(pts.shape[0]), _p(pts), _p(out))
    return out


def spec_expf(x) -> np.ndarray:
    """The blend's exponential (Cody-Waite + degree-7 Taylor, all fma)."""
    lib = _load()
    xa = _arr(x, np.float32).reshape(-1)
    y = np.zeros_like(xa)
    lib.orc_expf(ctypes.c_longlong(xa.size), _p(xa), _p(y))
    return y
