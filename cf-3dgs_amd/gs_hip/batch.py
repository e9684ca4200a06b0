"""Many pose solves at once: static workspaces, a captured HIP graph, streams.

The reference solves one frame at a time, each iteration a Python step with
a host synchronisation (the num_rendered readback of the rasteriser and
``torch.isnan(loss)``, trainer/losses.py:126).  Here a batch of independent
solves (one per frame instance) runs in lock-step:

  * every instance owns fixed buffers sized once (geometry, binning at a
    fixed capacity, image state, loss workspace, gradient scratch), so an
    iteration allocates nothing and never reads anything back;
  * one iteration of all instances -- preprocess (pose fused), binning +
    blend, fused L1+SSIM forward and backward (clamp to [0, 1] inside),
    render/preprocess backward asking for the pose gradient only -- is
    captured into a HIP graph, instance i on stream i mod S, joined before a
    single ``gs_se3_adam`` launch that updates every tangent and writes the
    next poses; ``iters_per_graph`` iterations are captured back to back and
    the graph is replayed until the iteration count is reached;
  * binning capacity is the number of tile-Gaussian pairs at the start pose
    times a headroom; an overflow raises a sticky device flag (the renderer
    then draws nothing for that iteration), read once after the solve, and
    the affected instances are re-solved with a larger capacity.

``autotune`` picks the stream count, the binning schedule, the blend /
render-backward schedule (per frame, or one launch over all instances, or
over each of 2, 4 or 8 instance groups, one stream per group) and whether the
depth sorts of a stream's or a group's instances run as one segmented sort,
per size bucket, by timing graph replays.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import _lib
from . import raster as _raster
from .loss import _window_taps
from .pose import EPS, ITERS, LR, PoseInstance, SE3Adam

_TAPS = None


def _taps():
    global _TAPS
    if _TAPS is None:
        _TAPS = _window_taps()
    return _TAPS


def _vp(t: Optional[torch.Tensor], offset_bytes: int = 0):
    return None if t is None else t.data_ptr() + offset_bytes


@dataclass
class _Slot:
    inst: PoseInstance
    capacity: int
    frame: _lib.GsFrame = None
    gauss: _lib.GsGaussians = None
    grads: _lib.GsGrads = None
    keep: list = field(default_factory=list)
    radii: torch.Tensor = None
    geom: torch.Tensor = None
    counters: torch.Tensor = None
    binning: torch.Tensor = None
    image: torch.Tensor = None
    color: torch.Tensor = None
    depth: torch.Tensor = None
    loss_ws: torch.Tensor = None
    dcolor: torch.Tensor = None
    scratch: torch.Tensor = None


def _round_capacity(n: int, headroom: float) -> int:
    q = 1 << 14
    c = int(n * headroom) + 4096
    return min(-(-c // q) * q, 0x7FFFFFFF)


def start_capacities(instances: Sequence[PoseInstance], poses: torch.Tensor, binning: int,
                     headroom: float) -> List[int]:
    """Binning capacity per instance: tile-Gaussian pairs at the given start
    poses ([n, 7] on the device) times ``headroom`` (one eager preprocess
    each, one readback for all)."""
    lib = _lib.load()
    n = len(instances)
    if n == 0:
        return []
    device = poses.device
    counts = torch.zeros((n, 4), dtype=torch.int32, device=device)
    stream = _raster._stream(device)
    keep = []
    for i, inst in enumerate(instances):
        fa = _raster.make_frame(inst.settings, device, poses[i], binning)
        ga = _raster.make_gaussians(inst.means, inst.shs, None, inst.opacities, inst.scales, inst.rotations,
                                    None, device)
        LG = _raster._geom_layout(inst.P)
        geom = torch.empty(max(LG.total, 256), dtype=torch.uint8, device=device)
        radii = torch.empty(inst.P, dtype=torch.int32, device=device)
        _lib.check(lib.gs_preprocess(ctypes.byref(fa.frame), ctypes.byref(ga.g), _raster._ptr(radii),
                                     _raster._ptr(geom), geom.numel(), _raster._ptr(counts[i]), stream),
                   "gs_preprocess")
        keep.append((fa, ga, geom, radii))
    host = counts.cpu()
    return [_round_capacity(int(host[i, 0]) & 0xFFFFFFFF, headroom) for i in range(n)]


def slot_bytes(P: int, W: int, H: int, capacity: int, binning: int) -> int:
    """Device bytes of one solve slot's buffers (the instance's own tensors excluded)."""
    lib = _lib.load()
    return int(_raster._geom_layout(P).total + _raster._binning_layout(capacity, P, W, H, binning).total
               + _raster._image_layout(W, H).total + 4 * P + 16 + 4 * 7 * H * W
               + lib.gs_loss_workspace_bytes(3, H, W) + lib.gs_backward_scratch_bytes(P, capacity, 1))


def make_slot(inst: PoseInstance, capacity: int, binning: int, pose_ptr: int, grad_ptr: int) -> _Slot:
    """Static buffers of one solve; the renderer reads the pose at
    ``pose_ptr`` (7 floats) and the backward writes dL/dpose to ``grad_ptr``."""
    lib = _lib.load()
    dev = inst.means.device
    s = _Slot(inst=inst, capacity=capacity)
    fa = _raster.make_frame(inst.settings, dev, None, binning, deterministic=True)
    fa.frame.pose = pose_ptr
    ga = _raster.make_gaussians(inst.means, inst.shs, None, inst.opacities, inst.scales, inst.rotations, None, dev)
    s.frame, s.gauss = fa.frame, ga.g
    s.keep = [fa, ga]
    P, W, H = inst.P, inst.W, inst.H
    LG = _raster._geom_layout(P)
    LI = _raster._image_layout(W, H)
    LB = _raster._binning_layout(capacity, P, W, H, binning)
    u8 = dict(dtype=torch.uint8, device=dev)
    s.geom = torch.empty(max(LG.total, 256), **u8)
    s.binning = torch.empty(max(LB.total, 256), **u8)
    s.image = torch.empty(LI.total, **u8)
    s.radii = torch.empty(P, dtype=torch.int32, device=dev)
    s.counters = torch.zeros(4, dtype=torch.int32, device=dev)
    s.color = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    s.depth = torch.empty((1, H, W), dtype=torch.float32, device=dev)
    s.dcolor = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    s.loss_ws = torch.empty(max(int(lib.gs_loss_workspace_bytes(3, H, W)), 256), **u8)
    s.scratch = torch.empty(max(int(lib.gs_backward_scratch_bytes(P, capacity, 1)), 256), **u8)
    g = _lib.GsGrads()
    g.dL_dpose = grad_ptr
    s.grads = g
    return s


def _bwd_args(s: _Slot):
    c = ctypes.c_void_p
    return (ctypes.byref(s.frame), ctypes.byref(s.gauss), c(s.radii.data_ptr()), c(s.geom.data_ptr()),
            c(s.binning.data_ptr()), s.capacity, c(s.image.data_ptr()), c(s.counters.data_ptr()),
            c(s.dcolor.data_ptr()), None, None, ctypes.byref(s.grads), c(s.scratch.data_ptr()), s.scratch.numel())


def launch_slot_binning(s: _Slot, stream: ctypes.c_void_p) -> None:
    """Preprocess (pose fused) and binning of one solve, enqueued on ``stream``."""
    lib = _lib.load()
    c = ctypes.c_void_p
    _lib.check(lib.gs_preprocess(ctypes.byref(s.frame), ctypes.byref(s.gauss), c(s.radii.data_ptr()),
                                 c(s.geom.data_ptr()), s.geom.numel(), c(s.counters.data_ptr()), stream),
               "gs_preprocess")
    _lib.check(lib.gs_render_binning(ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()),
                                     c(s.counters.data_ptr()), s.capacity, c(s.binning.data_ptr()),
                                     s.binning.numel(), c(s.image.data_ptr()), s.image.numel(), stream),
               "gs_render_binning")


def launch_slot_keys(s: _Slot, stream: ctypes.c_void_p) -> None:
    """Preprocess (pose fused) and the depth keys of one solve: the binning up
    to its depth sort, which a DepthSortSet then runs for many solves."""
    lib = _lib.load()
    c = ctypes.c_void_p
    _lib.check(lib.gs_preprocess(ctypes.byref(s.frame), ctypes.byref(s.gauss), c(s.radii.data_ptr()),
                                 c(s.geom.data_ptr()), s.geom.numel(), c(s.counters.data_ptr()), stream),
               "gs_preprocess")
    _lib.check(lib.gs_render_depth_keys(ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()),
                                        c(s.counters.data_ptr()), s.capacity, c(s.binning.data_ptr()),
                                        s.binning.numel(), c(s.image.data_ptr()), s.image.numel(), stream),
               "gs_render_depth_keys")


def launch_slot_after_sort(s: _Slot, stream: ctypes.c_void_p, blend: bool) -> None:
    """The binning after the depth sort (and the blend when ``blend``)."""
    lib = _lib.load()
    c = ctypes.c_void_p
    args = (ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()), c(s.counters.data_ptr()), s.capacity,
            c(s.binning.data_ptr()), s.binning.numel(), c(s.image.data_ptr()), s.image.numel())
    if blend:
        _lib.check(lib.gs_render_after_depth_sort(*args, c(s.color.data_ptr()), c(s.depth.data_ptr()), None, stream),
                   "gs_render_after_depth_sort")
    else:
        _lib.check(lib.gs_render_binning_after_depth_sort(*args, stream), "gs_render_binning_after_depth_sort")


def launch_slot_dup(s: _Slot, stream: ctypes.c_void_p) -> None:
    """After the depth sort: the tile-count prefix and the duplicate."""
    c = ctypes.c_void_p
    _lib.check(_lib.load().gs_render_binning_duplicate(ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()),
                                                       c(s.counters.data_ptr()), s.capacity, c(s.binning.data_ptr()),
                                                       s.binning.numel(), c(s.image.data_ptr()), s.image.numel(),
                                                       stream), "gs_render_binning_duplicate")


def launch_slot_tail(s: _Slot, stream: ctypes.c_void_p, blend: bool) -> None:
    """After the list sort: the checkpoint scan (and the blend when ``blend``)."""
    lib = _lib.load()
    c = ctypes.c_void_p
    args = (ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()), c(s.counters.data_ptr()), s.capacity,
            c(s.binning.data_ptr()), s.binning.numel(), c(s.image.data_ptr()), s.image.numel())
    if blend:
        _lib.check(lib.gs_render_after_list_sort(*args, c(s.color.data_ptr()), c(s.depth.data_ptr()), None, stream),
                   "gs_render_after_list_sort")
    else:
        _lib.check(lib.gs_render_binning_after_list_sort(*args, stream), "gs_render_binning_after_list_sort")


class ListSortSet:
    """The tile-list sorts of many solves as one segmented sort
    (gs_multi_list_sort).  Raises GsError for solves it cannot take
    (two-word list entries, a capacity above 2^21)."""

    def __init__(self, slots: Sequence[_Slot], device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        frames = (_lib.GsFrame * n)(*[s.frame for s in slots])
        P = (ctypes.c_int * n)(*[s.inst.P for s in slots])
        cap = (ctypes.c_longlong * n)(*[s.capacity for s in slots])
        bins = (vp * n)(*[s.binning.data_ptr() for s in slots])
        nbytes = (ctypes.c_size_t * n)(*[s.binning.numel() for s in slots])
        counters = (vp * n)(*[s.counters.data_ptr() for s in slots])
        images = (vp * n)(*[s.image.data_ptr() for s in slots])
        nb = int(lib.gs_multi_list_sort_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        self.blocks = (ctypes.c_uint * 4)()
        _lib.check(lib.gs_multi_list_sort_build(n, frames, P, cap, bins, nbytes, counters, images, host, nb,
                                                self.blocks), "gs_multi_list_sort_build")
        self.dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)

    def launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_list_sort(self.n, ctypes.c_void_p(self.dev.data_ptr()), self.blocks, stream),
                   "gs_multi_list_sort")


class LossSet:
    """The fused loss forward and backward of many solves in one launch of
    each kernel (gs_multi_loss); solve k writes its stats to out3 row k."""

    def __init__(self, slots: Sequence[_Slot], out3_ptrs: Sequence[int], one_ptr: int, lam: float, device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        arr = lambda vals: (vp * n)(*vals)  # noqa: E731
        ints = lambda vals: (ctypes.c_int * n)(*vals)  # noqa: E731
        nb = int(lib.gs_multi_loss_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        blocks = ctypes.c_uint(0)
        _lib.check(lib.gs_multi_loss_build(
            n, arr([s.color.data_ptr() for s in slots]), arr([s.inst.target.data_ptr() for s in slots]),
            ints([3] * n), ints([s.inst.H for s in slots]), ints([s.inst.W for s in slots]), _taps(), lam, 1,
            arr(list(out3_ptrs)), vp(one_ptr), arr([s.loss_ws.data_ptr() for s in slots]),
            (ctypes.c_size_t * n)(*[s.loss_ws.numel() for s in slots]), arr([s.dcolor.data_ptr() for s in slots]),
            host, nb, ctypes.byref(blocks)), "gs_multi_loss_build")
        self.dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.blocks = blocks.value

    def launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_loss(self.n, ctypes.c_void_p(self.dev.data_ptr()), self.blocks, stream),
                   "gs_multi_loss")


class DepthSortSet:
    """The depth sorts of many solves as one segmented sort
    (gs_multi_depth_sort): 12 launches for the set instead of 12 per solve."""

    def __init__(self, slots: Sequence[_Slot], device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        frames = (_lib.GsFrame * n)(*[s.frame for s in slots])
        P = (ctypes.c_int * n)(*[s.inst.P for s in slots])
        cap = (ctypes.c_longlong * n)(*[s.capacity for s in slots])
        bins = (vp * n)(*[s.binning.data_ptr() for s in slots])
        nbytes = (ctypes.c_size_t * n)(*[s.binning.numel() for s in slots])
        counters = (vp * n)(*[s.counters.data_ptr() for s in slots])
        nb = int(lib.gs_multi_depth_sort_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        blocks = ctypes.c_uint(0)
        _lib.check(lib.gs_multi_depth_sort_build(n, frames, P, cap, bins, nbytes, counters, host, nb,
                                                 ctypes.byref(blocks)), "gs_multi_depth_sort_build")
        self.dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.blocks = blocks.value

    def launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_depth_sort(self.n, ctypes.c_void_p(self.dev.data_ptr()), self.blocks, stream),
                   "gs_multi_depth_sort")


class KeysSet:
    """The preprocess and depth keys of many solves in three launches
    (gs_multi_preprocess_keys) instead of three per solve."""

    def __init__(self, slots: Sequence[_Slot], device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        sz = ctypes.c_size_t

        def ptrs(attr):
            return (vp * n)(*[getattr(s, attr).data_ptr() for s in slots])

        frames = (_lib.GsFrame * n)(*[s.frame for s in slots])
        gauss = (_lib.GsGaussians * n)(*[s.gauss for s in slots])
        cap = (ctypes.c_longlong * n)(*[s.capacity for s in slots])
        nb = int(lib.gs_multi_preprocess_keys_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        self.launch_dims = (ctypes.c_uint * 3)()
        _lib.check(lib.gs_multi_preprocess_keys_build(
            n, frames, gauss, ptrs("radii"), ptrs("geom"), (sz * n)(*[s.geom.numel() for s in slots]),
            ptrs("counters"), cap, ptrs("binning"), (sz * n)(*[s.binning.numel() for s in slots]), ptrs("image"),
            (sz * n)(*[s.image.numel() for s in slots]), host, nb, self.launch_dims),
            "gs_multi_preprocess_keys_build")
        self.dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)

    def launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_preprocess_keys(self.n, ctypes.c_void_p(self.dev.data_ptr()),
                                                        self.launch_dims, stream), "gs_multi_preprocess_keys")


class BinningStepsSet:
    """The binning steps on either side of the shared list sort of many
    solves: the duplicates (gs_multi_binning_duplicate, three launches) and
    the checkpoint scans (gs_multi_binning_after_list_sort, one launch)."""

    def __init__(self, slots: Sequence[_Slot], device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        sz = ctypes.c_size_t

        def ptrs(attr):
            return (vp * n)(*[getattr(s, attr).data_ptr() for s in slots])

        frames = (_lib.GsFrame * n)(*[s.frame for s in slots])
        P = (ctypes.c_int * n)(*[s.inst.P for s in slots])
        cap = (ctypes.c_longlong * n)(*[s.capacity for s in slots])
        nb = int(lib.gs_multi_binning_steps_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        blocks = ctypes.c_uint(0)
        _lib.check(lib.gs_multi_binning_steps_build(
            n, frames, P, ptrs("geom"), ptrs("counters"), cap, ptrs("binning"),
            (sz * n)(*[s.binning.numel() for s in slots]), ptrs("image"), (sz * n)(*[s.image.numel() for s in slots]),
            host, nb, ctypes.byref(blocks)), "gs_multi_binning_steps_build")
        self.dev = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.blocks = blocks.value

    def duplicate(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_binning_duplicate(self.n, ctypes.c_void_p(self.dev.data_ptr()), self.blocks,
                                                          stream), "gs_multi_binning_duplicate")

    def after_list_sort(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_binning_after_list_sort(self.n, ctypes.c_void_p(self.dev.data_ptr()), stream),
                   "gs_multi_binning_after_list_sort")


def launch_slot_loss(s: _Slot, out3_ptr: int, one_ptr: int, lam: float, stream: ctypes.c_void_p) -> None:
    """Fused loss forward (stats to ``out3_ptr``) and backward (d loss scaled
    by the float at ``one_ptr``) of one solve's rendered image."""
    lib = _lib.load()
    inst = s.inst
    c = ctypes.c_void_p
    _lib.check(lib.gs_ssim_l1_forward_backward(c(s.color.data_ptr()), c(inst.target.data_ptr()), 3, inst.H, inst.W,
                                               _taps(), lam, 1, c(out3_ptr), c(one_ptr), c(s.loss_ws.data_ptr()),
                                               s.loss_ws.numel(), c(s.dcolor.data_ptr()), stream),
               "gs_ssim_l1_forward_backward")


def launch_slot(s: _Slot, out3_ptr: int, one_ptr: int, lam: float, stream: ctypes.c_void_p) -> None:
    """One solve iteration minus the optimiser step, enqueued on ``stream``:
    preprocess (pose fused), binning + blend, fused loss forward and
    backward, pose-only backward with fixed-order reductions."""
    lib = _lib.load()
    inst = s.inst
    c = ctypes.c_void_p
    _lib.check(lib.gs_preprocess(ctypes.byref(s.frame), ctypes.byref(s.gauss), c(s.radii.data_ptr()),
                                 c(s.geom.data_ptr()), s.geom.numel(), c(s.counters.data_ptr()), stream),
               "gs_preprocess")
    _lib.check(lib.gs_render(ctypes.byref(s.frame), inst.P, c(s.geom.data_ptr()), c(s.counters.data_ptr()),
                             s.capacity, c(s.binning.data_ptr()), s.binning.numel(), c(s.image.data_ptr()),
                             s.image.numel(), c(s.color.data_ptr()), c(s.depth.data_ptr()), None, stream),
               "gs_render")
    launch_slot_loss(s, out3_ptr, one_ptr, lam, stream)
    _lib.check(lib.gs_backward(*_bwd_args(s), stream), "gs_backward")


def _lpt(items: Sequence[int], cost: Sequence[float], bins: int) -> List[List[int]]:
    """``items`` dealt longest first onto the least-loaded of ``bins`` bins
    (each bin's items in ascending order)."""
    load = [0.0] * bins
    out: List[List[int]] = [[] for _ in range(bins)]
    for i in sorted(items, key=lambda k: -cost[k]):
        j = min(range(bins), key=lambda q: load[q])
        out[j].append(i)
        load[j] += cost[i]
    return [sorted(b) for b in out]


def group_layout(cost: Sequence[float], nstreams: int, groups: int):
    """(group count, instances of each group, streams of each group, stream of
    each instance) for instances of the given costs.

    Every iteration ends when its slowest stream does, so work is dealt
    longest first onto the least-loaded bin: instances onto groups, then each
    group's instances onto its streams.  One group spreads its per-frame work
    over all ``nstreams`` streams.  Several groups take one stream each:
    every captured dependency must fork from or join into the capturing
    stream (HIP's end of capture fails when one side stream waits on
    another), so a group cannot fork streams of its own."""
    n = len(cost)
    G = max(1, min(int(groups), n))
    if G > 1 and G != nstreams:
        raise ValueError("with several instance groups each group runs on one stream: streams must equal groups")
    members = _lpt(range(n), cost, G)
    gstreams = [list(range(nstreams))] if G == 1 else [[j] for j in range(G)]
    stream_of = [0] * n
    for mem, sidx in zip(members, gstreams):
        for q, part in zip(sidx, _lpt(mem, cost, len(sidx))):
            for i in part:
                stream_of[i] = q
    return G, members, gstreams, stream_of


def schedule_candidates(binnings: Sequence[int], streams: Sequence[int], groups: Sequence[int],
                        shared: Sequence[bool] = (False, True)):
    """The autotuner's (binning, streams, groups, shared depth sort)
    candidates: per-frame (groups 0) and one group at every stream count,
    several groups only with as many streams; the shared depth sort only with
    the depth-first binning."""
    return [(b, s, g, ss) for b in binnings for s in streams for g in groups for ss in shared
            if (g <= 1 or g == s) and (not ss or b == _lib.BINNING_DEPTH_FIRST)]


class FrameSets:
    """Device descriptions of many solves' workspaces for the one-launch blend
    and render backward (gs_multi_blend / gs_multi_render_bwd) and, for
    pose-only frames, the backward's one-launch set-up and two-launch finish
    (gs_multi_pose_bwd_*)."""

    def __init__(self, slots: Sequence[_Slot], device):
        lib = _lib.load()
        n = len(slots)
        self.n = n
        vp = ctypes.c_void_p
        frames = (_lib.GsFrame * n)(*[s.frame for s in slots])
        P = (ctypes.c_int * n)(*[s.inst.P for s in slots])
        cap = (ctypes.c_longlong * n)(*[s.capacity for s in slots])

        def ptrs(attr):
            return (vp * n)(*[getattr(s, attr).data_ptr() for s in slots])

        nb = int(lib.gs_multi_blend_bytes(n))
        host = (ctypes.c_uint8 * nb)()
        tiles = ctypes.c_uint(0)
        _lib.check(lib.gs_multi_blend_build(n, frames, P, ptrs("geom"), cap, ptrs("binning"), ptrs("image"),
                                            ptrs("color"), ptrs("depth"), None, host, nb, ctypes.byref(tiles)),
                   "gs_multi_blend_build")
        self.blend = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.blend_tiles = tiles.value
        nr = int(lib.gs_multi_render_bwd_bytes(n))
        host = (ctypes.c_uint8 * nr)()
        blocks = ctypes.c_uint(0)
        _lib.check(lib.gs_multi_render_bwd_build(n, frames, P, ptrs("geom"), cap, ptrs("binning"), ptrs("image"),
                                                 ptrs("dcolor"), None, None, ptrs("scratch"), host, nr,
                                                 ctypes.byref(blocks)), "gs_multi_render_bwd_build")
        self.rbwd = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.rbwd_blocks = blocks.value
        # the pose solve's loss has no depth or alpha term: colour-only kernels,
        # and the pose-only backward when the library takes it for these frames
        self.rbwd_flags = int(slots[0].frame.deterministic) | _lib.GS_RB_COLOR_ONLY
        po = [lib.gs_backward_is_pose_only(ctypes.byref(s.frame), ctypes.byref(s.gauss), ctypes.byref(s.grads), 0)
              for s in slots]
        self.pose_bwd = None
        if all(po):
            self.rbwd_flags |= _lib.GS_RB_POSE_ONLY
            # and the pose-only set-up (one launch) and finish (per-Gaussian
            # pose partials, pose reduction: two launches) of all frames
            nf = int(lib.gs_multi_pose_bwd_bytes(n))
            host = (ctypes.c_uint8 * nf)()
            self.pose_bwd_blocks = (ctypes.c_uint * 2)()
            gauss = (_lib.GsGaussians * n)(*[s.gauss for s in slots])
            grads = (_lib.GsGrads * n)(*[s.grads for s in slots])
            sbytes = (ctypes.c_size_t * n)(*[s.scratch.numel() for s in slots])
            _lib.check(lib.gs_multi_pose_bwd_build(n, frames, gauss, ptrs("radii"), ptrs("geom"), ptrs("binning"),
                                                   cap, ptrs("image"), ptrs("counters"), ptrs("dcolor"), grads,
                                                   ptrs("scratch"), sbytes, host, nf, self.pose_bwd_blocks),
                       "gs_multi_pose_bwd_build")
            self.pose_bwd = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        elif any(po):
            raise RuntimeError("a render-backward set mixes pose-only and full frames")

    def blend_launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_blend(self.n, ctypes.c_void_p(self.blend.data_ptr()), self.blend_tiles,
                                              stream), "gs_multi_blend")

    def render_bwd_launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_render_bwd(self.n, ctypes.c_void_p(self.rbwd.data_ptr()),
                                                   self.rbwd_blocks, self.rbwd_flags, stream),
                   "gs_multi_render_bwd")

    def pose_bwd_prepare_launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_pose_bwd_prepare(self.n, ctypes.c_void_p(self.pose_bwd.data_ptr()),
                                                         self.pose_bwd_blocks, stream), "gs_multi_pose_bwd_prepare")

    def pose_bwd_finish_launch(self, stream: ctypes.c_void_p) -> None:
        _lib.check(_lib.load().gs_multi_pose_bwd_finish(self.n, ctypes.c_void_p(self.pose_bwd.data_ptr()),
                                                        self.pose_bwd_blocks, stream), "gs_multi_pose_bwd_finish")


class BatchedPoseSolver:
    """Solves ``len(instances)`` frames together.  ``run()`` returns
    (poses [n, 7], losses [n]) on the device."""

    def __init__(self, instances: Sequence[PoseInstance], iters: int = ITERS, lr: float = LR,
                 streams: int = 4, binning: int = _lib.BINNING_DEPTH_FIRST, iters_per_graph: int = 10,
                 headroom: float = 1.3, use_graph: bool = True, lambda_dssim: float = 0.2,
                 X0: Optional[torch.Tensor] = None, multi: bool = True, groups: int = 1,
                 shared_sort: bool = False):
        if not instances:
            raise ValueError("no instances")
        self.instances = list(instances)
        self.device = instances[0].means.device
        _raster.require_device(instances[0].means)
        self.n = len(self.instances)
        self.iters = int(iters)
        self.lr = float(lr)
        self.nstreams = max(1, int(streams))
        self.binning = int(binning)
        self.iters_per_graph = max(1, int(iters_per_graph))
        self.headroom = float(headroom)
        self.use_graph = bool(use_graph)
        self.lam = float(lambda_dssim)
        self.adam = SE3Adam(self.n, self.device, lr=lr, eps=EPS, X0=X0)
        self.one = torch.ones(1, dtype=torch.float32, device=self.device)
        self.out3 = torch.zeros((self.n, 3), dtype=torch.float32, device=self.device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.nstreams)]
        self.graph = None
        self.regrows = 0
        caps = start_capacities(self.instances, self.adam.pose, self.binning, self.headroom)
        self.slots = [self._make_slot(inst, cap, i) for i, (inst, cap) in enumerate(zip(self.instances, caps))]
        self.multi = bool(multi)
        # every iteration ends when its slowest stream does, so work is dealt
        # longest first onto the least-loaded bin (cost ~ list length + pixels):
        # instances onto groups (one-launch blend and backward per group; the
        # groups share no stream, so one group's launches overlap the others'
        # per-frame work), then each group's instances onto its streams
        cost = [cap + inst.W * inst.H for inst, cap in zip(self.instances, caps)]
        self.cost = cost
        self.ngroups, self.members, self.group_streams, self.stream_of = group_layout(
            cost, self.nstreams, int(groups) if self.multi else 1)
        self.sets = self._frame_sets() if self.multi else None
        # shared stages: the depth and list sorts as one segmented sort each
        # per group (one-launch blend schedules) or per stream (per-frame
        # schedules), and with one-launch blends the group's loss as one
        # launch of each loss kernel
        self.shared_sort = bool(shared_sort)
        if self.shared_sort and self.binning != _lib.BINNING_DEPTH_FIRST:
            raise ValueError("the shared depth sort needs the depth-first binning")
        self.stream_members = {q: [i for i in range(self.n) if self.stream_of[i] == q]
                               for q in sorted(set(self.stream_of))}
        self.dsets = self.lsets = self.ksets = self.msets = self.losssets = None
        if self.shared_sort:
            self.dsets, self.lsets, self.ksets, self.msets = self._sort_sets()
            self.losssets = self._loss_sets() if self.multi else None

    def _frame_sets(self) -> List[FrameSets]:
        # a one-launch stage dispatches its frames' workgroups frame after
        # frame: costliest frame first, so that its heavy tiles and work items
        # do not start last and run on alone (each frame's results do not
        # depend on the order)
        return [FrameSets([self.slots[i] for i in sorted(members, key=lambda i: -self.cost[i])], self.device)
                for members in self.members]

    def _loss_sets(self):
        return [LossSet([self.slots[i] for i in mem], [self.out3.data_ptr() + 12 * i for i in mem],
                        self.one.data_ptr(), self.lam, self.device) for mem in self.members]

    def _sort_sets(self):
        """Shared depth and list sorts, one per group (one-launch blend
        schedules) or per stream (per-frame schedules); the list sorts stay
        per frame (None) when a solve's list cannot join a shared one.  With
        them, the sets of the steps before and between the sorts (preprocess
        and depth keys; the duplicates and checkpoint scans, only beside
        shared list sorts)."""
        parts = self.members if self.multi else list(self.stream_members.values())
        keys = range(len(parts)) if self.multi else list(self.stream_members)
        try:
            dsets = {k: DepthSortSet([self.slots[i] for i in mem], self.device) for k, mem in zip(keys, parts)}
        except _lib.GsError as e:
            raise ValueError(f"the shared depth sort takes instances of at most 2^21 Gaussians ({e})") from e
        try:
            lsets = {k: ListSortSet([self.slots[i] for i in mem], self.device) for k, mem in zip(keys, parts)}
        except _lib.GsError:
            lsets = None
        ksets = {k: KeysSet([self.slots[i] for i in mem], self.device) for k, mem in zip(keys, parts)}
        msets = None if lsets is None else {k: BinningStepsSet([self.slots[i] for i in mem], self.device)
                                            for k, mem in zip(keys, parts)}
        return dsets, lsets, ksets, msets

    def _make_slot(self, inst: PoseInstance, capacity: int, i: int) -> _Slot:
        return make_slot(inst, capacity, self.binning, _vp(self.adam.pose, 28 * i), _vp(self.adam.grad, 28 * i))

    # -- one iteration ------------------------------------------------------
    def _launch_slot(self, s: _Slot, i: int, stream: ctypes.c_void_p):
        launch_slot(s, self.out3.data_ptr() + 12 * i, self.one.data_ptr(), self.lam, stream)

    def _per_slot(self, fn, members: Optional[Sequence[int]] = None, lead=None):
        """fn(i, slot, stream) for every slot (of ``members``), on its stream,
        between a fork from and a join into ``lead`` (default: the current
        stream)."""
        lead = torch.cuda.current_stream(self.device) if lead is None else lead
        members = range(self.n) if members is None else members
        used = sorted({self.stream_of[i] for i in members})
        for q in used:
            if self.streams[q] is not lead:
                self.streams[q].wait_stream(lead)
        for i in members:
            st = self.streams[self.stream_of[i]]
            with torch.cuda.stream(st):
                fn(i, self.slots[i], ctypes.c_void_p(st.cuda_stream))
        for q in used:
            if self.streams[q] is not lead:
                lead.wait_stream(self.streams[q])

    def _launch_iterations(self, k: int) -> None:
        for _ in range(k):
            self._launch_iteration()

    def _launch_iteration(self):
        if not self.multi and self.shared_sort:
            self._launch_iteration_shared_sort()
            return
        if not self.multi:
            self._per_slot(lambda i, s, st: self._launch_slot(s, i, st))
            self.adam.step()
            return
        # per group: per frame preprocess + binning; one blend launch over the
        # group; per frame loss and backward set-up; one render backward
        # launch; the per-Gaussian / pose gradients (for pose-only frames the
        # backward's set-up and these are one launch per kernel, else per
        # frame).  One group forks its per-frame
        # work from the current stream onto all streams and runs its
        # one-launch stages on the current stream; with more groups each
        # group runs everything on a stream of its own, forked from the
        # current stream, and the groups join before the step.  (Under graph
        # capture every dependency is a fork from or a join into the current
        # stream: HIP's end of capture fails when one side stream waits on
        # another.)
        lib = _lib.load()
        cur = torch.cuda.current_stream(self.device)

        def loss(i, s, st):
            launch_slot_loss(s, self.out3.data_ptr() + 12 * i, self.one.data_ptr(), self.lam, st)

        def loss_and_prepare(i, s, st):
            loss(i, s, st)
            _lib.check(lib.gs_backward_prepare(*_bwd_args(s), st), "gs_backward_prepare")

        def prepare(i, s, st):
            _lib.check(lib.gs_backward_prepare(*_bwd_args(s), st), "gs_backward_prepare")

        def finish(i, s, st):
            _lib.check(lib.gs_backward_finish(*_bwd_args(s), st), "gs_backward_finish")

        used = sorted(set(self.stream_of))
        if self.ngroups > 1:
            for q in used:
                self.streams[q].wait_stream(cur)
        for g, (members, sidx, sets) in enumerate(zip(self.members, self.group_streams, self.sets)):
            lead = cur if self.ngroups == 1 else self.streams[sidx[0]]
            lp = ctypes.c_void_p(lead.cuda_stream)
            if self.shared_sort:
                self.ksets[g].launch(lp)
                self.dsets[g].launch(lp)
                if self.lsets is not None:
                    self.msets[g].duplicate(lp)
                    self.lsets[g].launch(lp)
                    self.msets[g].after_list_sort(lp)
                else:
                    self._per_slot(lambda i, s, st: launch_slot_after_sort(s, st, blend=False), members, lead)
            else:
                self._per_slot(lambda i, s, st: launch_slot_binning(s, st), members, lead)
            sets.blend_launch(lp)
            pb = sets.pose_bwd is not None
            if self.losssets is not None:
                self.losssets[g].launch(lp)
                if not pb:
                    self._per_slot(prepare, members, lead)
            else:
                self._per_slot(loss if pb else loss_and_prepare, members, lead)
            if pb:
                sets.pose_bwd_prepare_launch(lp)
            sets.render_bwd_launch(lp)
            if pb:
                sets.pose_bwd_finish_launch(lp)
            else:
                self._per_slot(finish, members, lead)
        if self.ngroups > 1:
            for q in used:
                cur.wait_stream(self.streams[q])
        self.adam.step()

    def _launch_iteration_shared_sort(self):
        """Per-frame launches with each stream's depth sorts and list sorts as
        segmented sorts: a stream runs its solves' preprocess and depth keys,
        their shared depth sort, their duplicates, their shared list sort,
        then each solve's rest of the iteration."""
        lib = _lib.load()
        cur = torch.cuda.current_stream(self.device)
        for q in self.stream_members:
            self.streams[q].wait_stream(cur)
        for q, mem in self.stream_members.items():
            st = self.streams[q]
            sp = ctypes.c_void_p(st.cuda_stream)
            with torch.cuda.stream(st):
                self.ksets[q].launch(sp)
                self.dsets[q].launch(sp)
                if self.lsets is not None:
                    self.msets[q].duplicate(sp)
                    self.lsets[q].launch(sp)
                for i in mem:
                    s = self.slots[i]
                    if self.lsets is not None:
                        launch_slot_tail(s, sp, blend=True)
                    else:
                        launch_slot_after_sort(s, sp, blend=True)
                    launch_slot_loss(s, self.out3.data_ptr() + 12 * i, self.one.data_ptr(), self.lam, sp)
                    _lib.check(lib.gs_backward(*_bwd_args(s), sp), "gs_backward")
        for q in self.stream_members:
            cur.wait_stream(self.streams[q])
        self.adam.step()

    def _reset(self):
        self.adam.reset()
        for s in self.slots:
            s.counters.zero_()

    def _capture(self):
        # one eager iteration first (lazy loads, first-launch work); callers
        # reset the state before the captured iterations run
        self._launch_iteration()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_iterations(self.iters_per_graph)
        torch.cuda.synchronize(self.device)
        self.graph = g

    # -- solve ----------------------------------------------------------------
    def _run_once(self):
        self._reset()
        if self.use_graph:
            if self.graph is None:
                self._capture()
                self._reset()
            full, rest = divmod(self.iters, self.iters_per_graph)
            t0 = time.perf_counter()
            for _ in range(full):
                self.graph.replay()
            # host time inside the replays: a replay returns once its launches
            # are queued, so this approaches the wall time when submission,
            # not the GPU, sets the pace
            self.submit_seconds = time.perf_counter() - t0
            if rest:
                self._launch_iterations(rest)
        else:
            self._launch_iterations(self.iters)

    def overflowed(self) -> List[int]:
        flags = torch.stack([s.counters[3] for s in self.slots]).cpu()
        return [i for i in range(self.n) if int(flags[i]) & _lib.GS_STATUS_CAPACITY]

    def run(self, max_regrow: int = 4) -> Tuple[torch.Tensor, torch.Tensor]:
        for _ in range(max_regrow + 1):
            self._run_once()
            bad = self.overflowed()
            if not bad:
                return self.adam.pose.clone(), self.out3[:, 0].clone()
            # grow the instances that overflowed and solve the batch again
            self.regrows += 1
            self.graph = None
            for i in bad:
                s = self.slots[i]
                self.slots[i] = self._make_slot(s.inst, _round_capacity(int(s.capacity * 1.5), 1.0), i)
            if self.multi:
                self.sets = self._frame_sets()
            if self.shared_sort:
                self.dsets, self.lsets, self.ksets, self.msets = self._sort_sets()
                self.losssets = self._loss_sets() if self.multi else None
        raise RuntimeError(f"binning capacity still exceeded after {max_regrow} regrowths")

    def timed_run(self) -> Dict[str, float]:
        """run() with wall-clock seconds (graph capture excluded)."""
        if self.use_graph and self.graph is None:
            self._reset()
            self._capture()
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        poses, losses = self.run()
        torch.cuda.synchronize(self.device)
        return dict(seconds=time.perf_counter() - t0, poses=poses, losses=losses,
                    submit_seconds=getattr(self, "submit_seconds", None))


# ---------------------------------------------------------------------------
# autotuning of the batched solver per size bucket

_TUNED: Dict[tuple, dict] = {}


def size_bucket(instances: Sequence[PoseInstance]) -> tuple:
    """(instance count, total Gaussians rounded to 2^k, total pixels rounded to 2^k)."""
    P = sum(i.P for i in instances)
    px = sum(i.W * i.H for i in instances)
    return (len(instances), max(P, 1).bit_length(), max(px, 1).bit_length())


def autotune(instances: Sequence[PoseInstance], streams=(2, 4, 8),
             binnings=(_lib.BINNING_DEPTH_FIRST, _lib.BINNING_SORT64), groups=(0, 1, 2, 4, 8),
             shared=(False, True), iters_per_graph: int = 5, replays: int = 2) -> dict:
    """Times ``replays`` graph replays of every (streams, binning schedule,
    blend / render-backward schedule, shared depth sort) candidate and
    returns (and caches per size bucket) the fastest configuration.  The
    blend schedule is ``groups``: 0 launches the blend and the render
    backward per frame; g > 0 splits the instances into g groups with one
    launch of each over a group (one group spreads its per-frame work over
    all streams; several groups take one stream each, so g > 1 is tried with
    g streams).  The shared depth sort runs the depth sorts of a stream's
    (per frame) or a group's instances as one segmented sort."""
    key = size_bucket(instances)
    if key in _TUNED:
        return _TUNED[key]
    best = None
    results = []
    cands = schedule_candidates(binnings, streams, groups, shared)
    for b, s, g, ss in cands:
        try:
            solver = BatchedPoseSolver(instances, iters=iters_per_graph * replays, streams=s, binning=b,
                                       iters_per_graph=iters_per_graph, multi=g > 0, groups=max(g, 1),
                                       shared_sort=ss)
        except ValueError:  # a schedule these instances cannot take (the shared depth sort's size limit)
            continue
        solver._reset()
        solver._capture()
        solver._reset()
        solver.graph.replay()
        torch.cuda.synchronize(solver.device)
        t0 = time.perf_counter()
        for _ in range(replays):
            solver.graph.replay()
        torch.cuda.synchronize(solver.device)
        dt = (time.perf_counter() - t0) / (replays * iters_per_graph)
        results.append(dict(streams=s, binning=b, multi=g > 0, groups=g, shared_sort=ss, ms_per_iter=dt * 1e3))
        if best is None or dt < best[0]:
            best = (dt, s, b, g, ss)
        del solver
    cfg = dict(streams=best[1], binning=best[2], multi=best[3] > 0, groups=max(best[3], 1), shared_sort=best[4],
               ms_per_iter=best[0] * 1e3, candidates=results)
    _TUNED[key] = cfg
    return cfg
