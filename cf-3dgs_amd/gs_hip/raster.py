"""Torch-facing forward/backward of the gfx950 rasteriser.

Workspaces are torch uint8 tensors from the caching allocator; the library
itself never allocates.  ``num_rendered`` is read back once per forward after
preprocess, which is the reference's own synchronisation point (the binning
buffers are sized from it); everything else is asynchronous on the current
stream.
"""
from __future__ import annotations

import ctypes
import functools
import os
import threading
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib

# Binning schedule for every forward in this process: depth-first (default) or
# the reference's 64-bit key sort; both give identical sorted lists and ranges.
BINNING = (_lib.BINNING_SORT64 if os.environ.get("GS_BINNING", "").lower() == "sort64"
           else _lib.BINNING_DEPTH_FIRST)

# Backward reductions in a fixed order (bitwise reproducible gradients) for
# every backward in this process; off by default (float atomics, like the
# reference).  GS_DETERMINISTIC=1 or set_deterministic(True) turns it on.
DETERMINISTIC = os.environ.get("GS_DETERMINISTIC", "") not in ("", "0")


def set_deterministic(on: bool) -> None:
    global DETERMINISTIC
    DETERMINISTIC = bool(on)


def _dev_f32(t: Optional[torch.Tensor], device: torch.device, name: str) -> Optional[torch.Tensor]:
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t, dtype=torch.float32)
    if t.numel() == 0:
        return None
    if t.device.type == "cpu" and device.type == "cuda":
        t = t.to(device)
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if t.dtype != torch.float32:
        t = t.float()
    t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()  # kernels read [P,4] / [P,M,3] rows with 16-byte loads
    return t


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(t: torch.Tensor, name: str = "means3D") -> torch.device:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError(
            f"{name} must be a tensor on a HIP device: the rasteriser runs only on the GPU "
            "(libgs_hip.so, gfx950); there is no CPU path")
    return t.device


@dataclass
class FrameArgs:
    """Settings tensors kept alive for the duration of the launches."""
    frame: _lib.GsFrame
    view: torch.Tensor
    proj: torch.Tensor
    campos: Optional[torch.Tensor]
    bg: torch.Tensor
    pose: Optional[torch.Tensor]


def make_frame(settings, device: torch.device, pose: Optional[torch.Tensor] = None,
               binning: Optional[int] = None, deterministic: Optional[bool] = None) -> FrameArgs:
    view = _dev_f32(settings.viewmatrix, device, "viewmatrix")
    proj = _dev_f32(settings.projmatrix, device, "projmatrix")
    campos = _dev_f32(settings.campos, device, "campos")
    bg = _dev_f32(settings.bg, device, "bg")
    pose_t = _dev_f32(pose, device, "pose")
    if view is None or view.numel() != 16 or proj is None or proj.numel() != 16:
        raise ValueError("viewmatrix and projmatrix must hold 16 floats")
    if bg is None or bg.numel() != 3:
        raise ValueError("bg must hold 3 floats")
    if pose_t is not None and pose_t.numel() != 7:
        raise ValueError("pose must hold 7 floats [tx, ty, tz, qx, qy, qz, qw]")
    f = _lib.GsFrame()
    f.width = int(settings.image_width)
    f.height = int(settings.image_height)
    f.tanfovx = float(settings.tanfovx)
    f.tanfovy = float(settings.tanfovy)
    f.scale_modifier = float(settings.scale_modifier)
    f.sh_degree = int(settings.sh_degree)
    f.prefiltered = int(bool(settings.prefiltered))
    f.viewmatrix = view.data_ptr()
    f.projmatrix = proj.data_ptr()
    f.campos = campos.data_ptr() if campos is not None else None
    f.bg = bg.data_ptr()
    f.pose = pose_t.data_ptr() if pose_t is not None else None
    f.binning = BINNING if binning is None else int(binning)
    f.deterministic = int(DETERMINISTIC if deterministic is None else bool(deterministic))
    return FrameArgs(f, view, proj, campos, bg, pose_t)


@dataclass
class GaussianArgs:
    g: _lib.GsGaussians
    tensors: tuple


def make_gaussians(means3D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp,
                   device: torch.device) -> GaussianArgs:
    P = means3D.shape[0]
    means = _dev_f32(means3D, device, "means3D")
    if means is None:
        means = torch.zeros((0, 3), device=device)
    if means.dim() != 2 or means.shape[1] != 3:
        raise ValueError("means3D must be [P, 3]")
    sh = _dev_f32(shs, device, "shs")
    col = _dev_f32(colors_precomp, device, "colors_precomp")
    op = _dev_f32(opacities, device, "opacities")
    sc = _dev_f32(scales, device, "scales")
    ro = _dev_f32(rotations, device, "rotations")
    cv = _dev_f32(cov3D_precomp, device, "cov3D_precomp")
    M = 0
    if sh is not None:
        if sh.dim() != 3 or sh.shape[0] != P or sh.shape[2] != 3:
            raise ValueError("shs must be [P, M, 3]")
        M = sh.shape[1]
    if col is not None and col.numel() != 3 * P:
        raise ValueError("colors_precomp must be [P, 3]")
    if (op is None and P > 0) or (op is not None and op.numel() != P):
        raise ValueError("opacities must hold P values")
    if sc is not None and sc.numel() != 3 * P:
        raise ValueError("scales must be [P, 3]")
    if ro is not None and ro.numel() != 4 * P:
        raise ValueError("rotations must be [P, 4]")
    if cv is not None and cv.numel() != 6 * P:
        raise ValueError("cov3D_precomp must be [P, 6]")
    g = _lib.GsGaussians()
    g.P = P
    g.M = M
    g.means3D = means.data_ptr()
    g.shs = sh.data_ptr() if sh is not None else None
    g.colors_precomp = col.data_ptr() if col is not None else None
    g.opacities = op.data_ptr() if op is not None else None
    g.scales = sc.data_ptr() if sc is not None else None
    g.rotations = ro.data_ptr() if ro is not None else None
    g.cov3D_precomp = cv.data_ptr() if cv is not None else None
    return GaussianArgs(g, (means, sh, col, op, sc, ro, cv))


_CAPACITY_QUANTUM = 1 << 16
_CAPACITY_MAX = 0x7FFFFFFF  # list entries are indexed with 32-bit offsets
# per thread: a pinned counter buffer and an event per device for the
# num_rendered readback (a forward waits for its own copy before returning,
# so calls on one thread never overlap; calls on different threads get
# different buffers)
_TLS = threading.local()
# capacity of the last forward per (P, W, H, schedule): the guess under which
# the part of the binning that does not need num_rendered is queued before the
# readback
_LAST_CAPACITY = {}


def _readback(device: torch.device):
    """(pinned int32[4], event) of the calling thread for ``device``."""
    d = getattr(_TLS, "readback", None)
    if d is None:
        d = _TLS.readback = {}
    rb = d.get(device)
    if rb is None:
        rb = d[device] = (torch.empty(4, dtype=torch.int32, pin_memory=True), torch.cuda.Event())
    return rb


@functools.lru_cache(maxsize=256)
def _geom_layout(P: int):
    return _lib.geom_layout(P)


@functools.lru_cache(maxsize=64)
def _image_layout(W: int, H: int):
    return _lib.image_layout(W, H)


@functools.lru_cache(maxsize=256)
def _binning_layout(capacity: int, P: int, W: int, H: int, mode: int):
    return _lib.binning_layout(capacity, P, W, H, mode)


@dataclass
class ForwardState:
    """Buffers the backward needs (the reference keeps geomBuffer,
    binningBuffer and imgBuffer alive in the autograd context)."""
    geom: torch.Tensor
    binning: torch.Tensor
    image: torch.Tensor
    counters: torch.Tensor
    radii: torch.Tensor
    capacity: int
    num_rendered: int
    binning_mode: int = 0
    frame_args: Optional["FrameArgs"] = None        # the forward's, reused by the backward
    gaussian_args: Optional["GaussianArgs"] = None


def _same_tensors(held, given) -> bool:
    """Whether the tensors an argument struct holds are the given ones (absent
    or empty on both sides, or the same storage)."""
    for h, g in zip(held, given):
        if g is None or (isinstance(g, torch.Tensor) and g.numel() == 0):
            if h is not None:
                return False
        elif h is None or not isinstance(g, torch.Tensor) or h.data_ptr() != g.data_ptr() or h.device != g.device:
            return False
    return True


def forward(settings, means3D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp,
            pose=None, want_alpha: bool = True, binning: Optional[int] = None):
    """Returns (color [3,H,W], radii [P] int32, depth [1,H,W], alpha [1,H,W] or None, state)."""
    lib = _lib.load()
    device = require_device(means3D)
    fa = make_frame(settings, device, pose, binning)
    ga = make_gaussians(means3D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp, device)
    P = ga.g.P
    W, H = fa.frame.width, fa.frame.height
    stream = _stream(device)
    LG = _geom_layout(P)
    LI = _image_layout(W, H)
    geom = torch.empty(max(LG.total, 256), dtype=torch.uint8, device=device)
    radii = torch.empty(P, dtype=torch.int32, device=device)
    counters = torch.empty(4, dtype=torch.int32, device=device)
    # everything not sized by num_rendered is allocated before the readback
    image = torch.empty(LI.total, dtype=torch.uint8, device=device)
    color = torch.empty((3, H, W), dtype=torch.float32, device=device)
    depth = torch.empty((1, H, W), dtype=torch.float32, device=device)
    alpha = torch.empty((1, H, W), dtype=torch.float32, device=device) if want_alpha else None
    _lib.check(lib.gs_preprocess(ctypes.byref(fa.frame), ctypes.byref(ga.g), _ptr(radii), _ptr(geom),
                                 geom.numel(), _ptr(counters), stream), "gs_preprocess")
    # the one host synchronisation of the forward (the reference's num_rendered
    # readback): only the copy is waited for.  Meanwhile the binning work that
    # does not need the count runs under the last capacity used for this size.
    host, ev = _readback(device)
    host.copy_(counters, non_blocking=True)
    ev.record(torch.cuda.current_stream(device))
    key = (P, W, H, fa.frame.binning)
    guess = _LAST_CAPACITY.get(key, 0)
    binning = None
    if guess:
        LB = _binning_layout(guess, P, W, H, fa.frame.binning)
        binning = torch.empty(max(LB.total, 256), dtype=torch.uint8, device=device)
        _lib.check(lib.gs_render_prepare(ctypes.byref(fa.frame), P, _ptr(geom), _ptr(counters), guess,
                                         _ptr(binning), binning.numel(), _ptr(image), image.numel(), stream),
                   "gs_render_prepare")
    ev.synchronize()
    num_rendered = int(host[0]) & 0xFFFFFFFF
    status = int(host[1])
    if status & _lib.GS_STATUS_PREFILTERED:
        raise RuntimeError("a point was culled by the near plane although prefiltered=True")
    # capacity rounded up so the caching allocator can reuse the binning block
    if num_rendered > _CAPACITY_MAX:
        raise RuntimeError(f"{num_rendered} list entries exceed the binning limit of {_CAPACITY_MAX}")
    capacity = min(-(-num_rendered // _CAPACITY_QUANTUM) * _CAPACITY_QUANTUM, _CAPACITY_MAX)
    if guess and num_rendered <= guess and capacity * 2 > guess:
        capacity = guess
        _lib.check(lib.gs_render_resume(ctypes.byref(fa.frame), P, _ptr(geom), _ptr(counters), capacity,
                                        _ptr(binning), binning.numel(), _ptr(image), image.numel(), _ptr(color),
                                        _ptr(depth), _ptr(alpha), stream), "gs_render_resume")
    else:  # first frame of this size, or the list grew past (or shrank well below) the guess
        LB = _binning_layout(capacity, P, W, H, fa.frame.binning)
        binning = torch.empty(max(LB.total, 256), dtype=torch.uint8, device=device)
        _lib.check(lib.gs_render(ctypes.byref(fa.frame), P, _ptr(geom), _ptr(counters), capacity, _ptr(binning),
                                 binning.numel(), _ptr(image), image.numel(), _ptr(color), _ptr(depth), _ptr(alpha),
                                 stream), "gs_render")
    _LAST_CAPACITY[key] = capacity
    st = ForwardState(geom, binning, image, counters, radii, capacity, num_rendered, fa.frame.binning, fa, ga)
    return color, radii, depth, alpha, st


def backward(settings, st: ForwardState, means3D, shs, colors_precomp, opacities, scales, rotations,
             cov3D_precomp, pose, grad_color, grad_depth=None, grad_alpha=None, want=None,
             deterministic: Optional[bool] = None, adam=None):
    """Returns a dict of gradients (None where the input was absent or, when
    `want` (a set of names) is given, not wanted).  With ``adam`` (a
    gs_adam_fused_t, ``FusedAdam.in_backward(...).take()``, or a dict of its
    fields) the
    means and SH coefficients take one Adam step inside the backward
    (gs_backward_adam, updated in place) and get no gradient."""
    lib = _lib.load()
    device = require_device(means3D)
    # the forward's argument structs when they describe the same call (the
    # autograd function passes the tensors it saved, so the pointers match)
    want_det = int(DETERMINISTIC if deterministic is None else bool(deterministic))
    fa = st.frame_args
    if fa is None or fa.frame.deterministic != want_det or not _same_tensors((fa.pose,), (pose,)):
        fa = make_frame(settings, device, pose, st.binning_mode, deterministic)
    ga = st.gaussian_args
    if ga is None or not _same_tensors(ga.tensors, (means3D, shs, colors_precomp, opacities, scales, rotations,
                                                    cov3D_precomp)):
        ga = make_gaussians(means3D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp, device)
    P = ga.g.P
    M = ga.g.M
    stream = _stream(device)
    if adam is not None:  # these two take their step in place and get no gradient array
        every = {"means2D", "colors", "opacities", "means3D", "cov3D", "shs", "scales", "rotations", "pose"}
        want = (set(want) if want is not None else every) - {"means3D", "shs"}
    dc = _dev_f32(grad_color, device, "grad_color")
    dd = _dev_f32(grad_depth, device, "grad_depth")
    da = _dev_f32(grad_alpha, device, "grad_alpha")
    if dc is None:
        dc = torch.zeros((3, fa.frame.height, fa.frame.width), device=device)
    f32 = dict(dtype=torch.float32, device=device)

    def alloc(name, shape, present=True):
        if not present or (want is not None and name not in want):
            return None
        return torch.empty(shape, **f32)

    out = dict(
        means2D=alloc("means2D", (P, 3)),
        colors=alloc("colors", (P, 3)),
        opacities=alloc("opacities", (P,)),
        means3D=alloc("means3D", (P, 3)),
        cov3D=alloc("cov3D", (P, 6)),
        shs=alloc("shs", (P, M, 3), bool(ga.g.shs)),
        scales=alloc("scales", (P, 3), bool(ga.g.scales)),
        rotations=alloc("rotations", (P, 4), bool(ga.g.rotations)),
        pose=alloc("pose", (7,), fa.pose is not None),
    )

    def ptr(name):
        t = out[name]
        return t.data_ptr() if t is not None else None

    gr = _lib.GsGrads()
    gr.dL_dmeans2D = ptr("means2D")
    gr.dL_dcolors = ptr("colors")
    gr.dL_dopacity = ptr("opacities")
    gr.dL_dmeans3D = ptr("means3D")
    gr.dL_dcov3D = ptr("cov3D")
    gr.dL_dsh = ptr("shs")
    gr.dL_dscales = ptr("scales")
    gr.dL_drotations = ptr("rotations")
    gr.dL_dpose = ptr("pose")
    scratch = torch.empty(lib.gs_backward_scratch_bytes(P, st.capacity, fa.frame.deterministic), dtype=torch.uint8,
                          device=device)
    if adam is None:
        _lib.check(lib.gs_backward(ctypes.byref(fa.frame), ctypes.byref(ga.g), _ptr(st.radii), _ptr(st.geom),
                                   _ptr(st.binning), st.capacity, _ptr(st.image), _ptr(st.counters), _ptr(dc),
                                   _ptr(dd), _ptr(da), ctypes.byref(gr), _ptr(scratch), scratch.numel(), stream),
                   "gs_backward")
        return out
    # the parameters are updated in place: the kernel must see the caller's
    # storage, not a converted copy
    for name, t, k in (("means3D", means3D, 0), ("shs", shs, 1)):
        if t is not None and (ga.tensors[k] is None or ga.tensors[k].data_ptr() != t.data_ptr()):
            raise RuntimeError(f"adam in backward: {name} must be a contiguous float32 device tensor")
    ad = adam if isinstance(adam, _lib.GsAdamFused) else _lib.GsAdamFused(**adam)
    _lib.check(lib.gs_backward_adam(ctypes.byref(fa.frame), ctypes.byref(ga.g), _ptr(st.radii), _ptr(st.geom),
                                    _ptr(st.binning), st.capacity, _ptr(st.image), _ptr(st.counters), _ptr(dc),
                                    _ptr(dd), _ptr(da), ctypes.byref(gr), ctypes.byref(ad), _ptr(scratch),
                                    scratch.numel(), stream), "gs_backward_adam")
    return out


def mark_visible(positions: torch.Tensor, viewmatrix, projmatrix) -> torch.Tensor:
    lib = _lib.load()
    device = require_device(positions, "positions")
    pos = _dev_f32(positions, device, "positions")
    view = _dev_f32(viewmatrix, device, "viewmatrix")
    proj = _dev_f32(projmatrix, device, "projmatrix")
    P = positions.shape[0]
    vis = torch.empty(P, dtype=torch.uint8, device=device)
    if P:
        _lib.check(lib.gs_mark_visible(_ptr(pos), P, _ptr(view), _ptr(proj), _ptr(vis), _stream(device)),
                   "gs_mark_visible")
    return vis.bool()


def knn3_meandist(points: torch.Tensor) -> torch.Tensor:
    lib = _lib.load()
    device = require_device(points, "points")
    pts = _dev_f32(points, device, "points")
    if pts is None:
        return torch.empty(0, dtype=torch.float32, device=device)
    if pts.dim() != 2 or pts.shape[1] != 3:
        raise ValueError("points must be [P, 3]")
    P = pts.shape[0]
    out = torch.empty(P, dtype=torch.float32, device=device)
    ws = torch.empty(max(int(lib.gs_knn_workspace_bytes(P)), 256), dtype=torch.uint8, device=device)
    _lib.check(lib.gs_knn3_meandist(_ptr(pts), P, _ptr(out), _ptr(ws), ws.numel(), _stream(device)),
               "gs_knn3_meandist")
    return out


def sort_pairs_u64(keys: torch.Tensor, vals: torch.Tensor, end_bit: int):
    """Stable radix sort of int64 (bit pattern of uint64) keys with int32 values (test entry)."""
    lib = _lib.load()
    device = require_device(keys, "keys")
    n = keys.numel()
    ka = keys.contiguous().clone()
    va = vals.to(torch.int32).contiguous().clone()
    kb = torch.empty_like(ka)
    vb = torch.empty_like(va)
    n_dev = torch.tensor([n], dtype=torch.int32, device=device)
    tmp = torch.empty(max(int(lib.gs_sort_tmp_bytes(n)), 256), dtype=torch.uint8, device=device)
    in_b = ctypes.c_int(0)
    _lib.check(lib.gs_sort_pairs_u64(_ptr(ka), _ptr(va), _ptr(kb), _ptr(vb), _ptr(n_dev), n, end_bit, _ptr(tmp),
                                     tmp.numel(), ctypes.byref(in_b), _stream(device)), "gs_sort_pairs_u64")
    return (kb, vb) if in_b.value else (ka, va)


def views(st: ForwardState, P: int, W: int, H: int) -> dict:
    """Typed views of the internal buffers (keys, sorted ids, ranges, ...) for parity checks."""
    LG = _lib.geom_layout(P)
    LB = _lib.binning_layout(st.capacity, P, W, H, st.binning_mode)
    LI = _lib.image_layout(W, H)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    n = st.num_rendered if st.num_rendered <= st.capacity else 0

    def sl(buf, off, nbytes, dtype):
        return buf[off:off + nbytes].view(dtype)

    splat = sl(st.geom, LG.splat, 48 * P, torch.float32).view(P, 3, 4)
    if LB.id_bits:  # one word per entry: tile << id_bits | id
        words = sl(st.binning, LB.keys_sorted, 4 * n, torch.int32).to(torch.int64) & 0xFFFFFFFF
        vals = (words & ((1 << LB.id_bits) - 1)).to(torch.int32)
        tiles_words = words >> LB.id_bits
    else:
        vals = sl(st.binning, LB.vals_sorted, 4 * n, torch.int32)
    if LB.key_bytes == 8:
        keys = sl(st.binning, LB.keys_sorted, 8 * n, torch.int64)
    else:
        # (tile << 32 | depth bits) rebuilt from the sorted tile ids and the ids' depths
        tiles_sorted = (tiles_words if LB.id_bits else
                        sl(st.binning, LB.keys_sorted, 4 * n, torch.int32).to(torch.int64))
        dbits = splat[:, 1, 2].contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        keys = (tiles_sorted << 32) | dbits[vals.to(torch.int64)] if n else torch.zeros(0, dtype=torch.int64)
    return dict(
        xy=splat[:, 0, 0:2],
        conic=splat[:, 0, 2:4],
        conic_c=splat[:, 1, 0],
        opacity=splat[:, 1, 1],
        depth=splat[:, 1, 2],
        clamp_bits=splat[:, 1, 3].view(torch.int32) if P else torch.zeros(0, dtype=torch.int32),
        rgb=splat[:, 2, 0:3],
        cov3D=sl(st.geom, LG.cov3D, 24 * P, torch.float32).view(P, 6),
        tiles=sl(st.geom, LG.tiles, 4 * P, torch.int32),
        keys=keys,
        vals=vals,
        ranges=sl(st.image, LI.ranges, 8 * gx * gy, torch.int32).view(gx * gy, 2),
        final_T=sl(st.image, LI.final_T, 4 * W * H, torch.float32).view(H, W),
        n_contrib=sl(st.image, LI.n_contrib, 4 * W * H, torch.int32).view(H, W),
    )
