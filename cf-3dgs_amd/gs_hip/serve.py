"""Serving loop: resident scenes, render and pose-solve requests, continuous batching.

A request is either
  * a render of a resident scene at a camera pose (forward only), or
  * a pose solve of a new frame: the frame-ingest of trainer/trainer.py:476-512
    (depth lift, distCUDA2, create_from_pcd), done on the device from the
    request's depth map and image, followed by the per-frame solve of
    trainer/cf3dgs_trainer.py:261-285 (``iters`` iterations of render -> loss
    -> backward -> Adam on the SE(3) tangent).

The host loop never waits for the whole device.  Each request runs on its
own stream (round-robin over ``streams``) and reports progress through
events:
  * arrived requests are queued and admitted in arrival order while the
    memory budget and the slot count allow (a request at the head that does
    not fit blocks the ones behind it and is counted as having waited for
    memory; one that could never fit is rejected);
  * a render queues its preprocess and an asynchronous copy of its list
    length; once that copy has landed, its binning and blend; once those
    are done it is retired and its memory released;
  * a solve keeps up to two chunks of ``iters_per_tick`` iterations queued
    on its stream and queues the next chunk as soon as the older one has
    finished;
so requests that arrive while others are on the device are admitted as
memory frees up and their kernels overlap with the running ones.

Solves use the fixed-order backward, and renders are deterministic, so a
result does not depend on what else was in flight: a request served alone
and the same request served in a mix give identical bits.
"""
from __future__ import annotations

import ctypes
import math
import random
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import raster as _raster
from .batch import _Slot, launch_slot, make_slot, slot_bytes, start_capacities
from .camera import focal2fov, fov2focal, look_at, make_camera, raster_settings
from .pose import EPS, LR, PoseInstance, SE3Adam, build_instance, build_instance_device

GB = 1 << 30


@dataclass
class Scene:
    """Activated Gaussian tensors of one scene, resident on the device."""
    sid: int
    means: torch.Tensor      # [P, 3]
    shs: torch.Tensor        # [P, 16, 3]
    opacities: torch.Tensor  # [P, 1]
    scales: torch.Tensor     # [P, 3]
    rotations: torch.Tensor  # [P, 4]
    sh_degree: int = 3

    @property
    def P(self) -> int:
        return int(self.means.shape[0])

    @property
    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in
                   (self.means, self.shs, self.opacities, self.scales, self.rotations))


def make_scene(sid: int, P: int, device, seed: Optional[int] = None, n_clusters: int = 32) -> Scene:
    """A seeded scene like the training workload: means in isotropic clusters
    (sigma 0.3) with centres U([-3, 3]^3), log-scales N(-3, 0.5^2),
    normalised N(0, I4) rotations, sigmoid(N(0, 1)) opacities, SH degree 3
    coefficients N(0, 0.1^2); activations applied as get_scaling /
    get_rotation / get_opacity do (scene/gaussian_model_cf.py:115-121,157-169)."""
    g = torch.Generator().manual_seed(1000 + sid if seed is None else seed)
    centres = torch.rand((n_clusters, 3), generator=g) * 6 - 3
    assign = torch.randint(0, n_clusters, (P,), generator=g)
    xyz = centres[assign] + torch.randn((P, 3), generator=g) * 0.3
    scales = torch.exp(torch.randn((P, 3), generator=g) * 0.5 - 3.0)
    rot = torch.nn.functional.normalize(torch.randn((P, 4), generator=g), dim=-1)
    opac = torch.sigmoid(torch.randn((P, 1), generator=g))
    shs = torch.randn((P, 16, 3), generator=g) * 0.1
    dev = torch.device(device)
    return Scene(sid, xyz.to(dev), shs.to(dev), opac.to(dev), scales.to(dev), rot.to(dev))


def sphere_camera(seed: int, W: int, H: int, radius: float = 4.0, fovx_deg: float = 60.0, device="cuda"):
    """Settings for a camera at a seeded random point of the sphere of
    ``radius`` around the origin, looking at the origin."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=3)
    eye = radius * v / np.linalg.norm(v)
    R, t = look_at(eye, np.zeros(3))
    fovx = math.radians(fovx_deg)
    cam = make_camera(R, t, fovx, focal2fov(fov2focal(fovx, W), H), W, H).to(torch.device(device))
    return cam


@dataclass
class Request:
    rid: int
    kind: str                 # "render" | "solve"
    arrival: float            # seconds from the start of the run
    scene: int = 0            # render: scene id
    W: int = 480
    H: int = 270
    seed: int = 0             # render: camera seed; solve: frame seed
    P: Optional[int] = None   # solve: Gaussians of the new frame (None = one per pixel)
    # filled by the server
    admitted: Optional[float] = None
    done: Optional[float] = None
    waited_for_memory: bool = False
    reserved: int = 0
    result: Optional[torch.Tensor] = None
    loss: Optional[float] = None
    error: Optional[str] = None   # set when the request can never fit the budget

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Render:
    req: Request
    stream: torch.cuda.Stream
    frame: object = None
    gauss: object = None
    keep: list = field(default_factory=list)
    geom: torch.Tensor = None
    radii: torch.Tensor = None
    counters: torch.Tensor = None
    image: torch.Tensor = None
    binning: torch.Tensor = None
    color: torch.Tensor = None
    depth: torch.Tensor = None
    event: object = None       # the list-length copy, then the end of the blend
    host: torch.Tensor = None  # pinned copy of the counters
    launched: bool = False     # binning + blend queued


@dataclass
class _Solve:
    req: Request
    stream: torch.cuda.Stream
    inst: PoseInstance
    slot: _Slot
    adam: SE3Adam
    out3: torch.Tensor
    iters_done: int = 0
    chunks: deque = field(default_factory=deque)  # events of the queued iteration chunks
    host: torch.Tensor = None   # pinned copy of the sticky status word at the end
    final: object = None        # its event


class Server:
    """The request loop (see the module docstring)."""

    def __init__(self, scenes: Sequence[Scene], budget_bytes: float = 200 * GB, max_slots: int = 32,
                 streams: int = 4, solve_iters: int = 200, iters_per_tick: int = 10, lr: float = LR,
                 binning: int = _lib.BINNING_DEPTH_FIRST, headroom: float = 1.3, render_dup_estimate: float = 8.0,
                 seen_counts: Optional[Dict[tuple, int]] = None):
        if not scenes:
            raise ValueError("no scenes")
        self.scenes = {s.sid: s for s in scenes}
        self.device = scenes[0].means.device
        _raster.require_device(scenes[0].means)
        self.budget = float(budget_bytes)
        self.resident = sum(s.nbytes for s in scenes)
        if self.resident > self.budget:
            raise ValueError("the resident scenes alone exceed the memory budget")
        self.max_slots = int(max_slots)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(streams)))]
        self.ingest_stream = torch.cuda.Stream(device=self.device)
        self.solve_iters = int(solve_iters)
        self.iters_per_tick = max(1, int(iters_per_tick))
        self.lr = float(lr)
        self.binning = int(binning)
        self.headroom = float(headroom)
        self.dup_est = float(render_dup_estimate)
        self.bg = torch.zeros(3, device=self.device)
        self.one = torch.ones(1, dtype=torch.float32, device=self.device)
        self._next_stream = 0
        self.used = float(self.resident)
        self.peak_used = float(self.used)
        # (scene, W, H) -> last list length; a server may start from another's
        self.seen_counts: Dict[tuple, int] = dict(seen_counts or {})
        self.max_request_bytes = 0  # largest final reservation of a served request
        self.ticks = 0

    # -- memory estimates ------------------------------------------------------
    @staticmethod
    def _round_cap(n: int) -> int:
        return min(-(-max(int(n), 1) // 4096) * 4096, 0x7FFFFFFF)

    def _render_bytes(self, req: Request, capacity: Optional[int] = None) -> int:
        """Device bytes of a render: its workspaces at ``capacity`` list
        entries, or, before the count is known, at the last count seen for
        the same scene and resolution (+5%), else ``render_dup_estimate``
        entries per Gaussian."""
        sc = self.scenes[req.scene]
        P = sc.P
        if capacity is not None:
            cap = capacity
        else:
            seen = self.seen_counts.get((req.scene, req.W, req.H))
            cap = self._round_cap(seen * 1.05) if seen is not None else int(P * self.dup_est)
        return int(_raster._geom_layout(P).total + _raster._image_layout(req.W, req.H).total
                   + _raster._binning_layout(min(cap, 0x7FFFFFFF), P, req.W, req.H, self.binning).total
                   + 4 * P + 16 + 4 * 4 * req.W * req.H)

    def _solve_bytes(self, req: Request) -> int:
        P = req.P if req.P is not None else req.W * req.H
        per_gauss = 4 * (3 + 3 + 1 + 3 + 4)  # the frame's own Gaussians
        cap = int(P * 4 * self.headroom) + 4096  # start estimate; corrected after the capacity readback
        return int(per_gauss * P + 4 * 3 * req.W * req.H + slot_bytes(P, req.W, req.H, cap, self.binning))

    def _stream(self) -> torch.cuda.Stream:
        st = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        return st

    # -- request set-up ----------------------------------------------------------
    def _start_render(self, req: Request) -> _Render:
        lib = _lib.load()
        sc = self.scenes[req.scene]
        r = _Render(req=req, stream=self._stream())
        cam = sphere_camera(req.seed, req.W, req.H, device=self.device)
        rs = raster_settings(cam, self.bg, sc.sh_degree)
        fa = _raster.make_frame(rs, self.device, None, self.binning)
        ga = _raster.make_gaussians(sc.means, sc.shs, None, sc.opacities, sc.scales, sc.rotations, None, self.device)
        r.frame, r.gauss, r.keep = fa.frame, ga.g, [fa, ga, rs]
        LG = _raster._geom_layout(sc.P)
        LI = _raster._image_layout(req.W, req.H)
        u8 = dict(dtype=torch.uint8, device=self.device)
        r.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(r.stream):
            r.geom = torch.empty(max(LG.total, 256), **u8)
            r.image = torch.empty(LI.total, **u8)
            r.radii = torch.empty(sc.P, dtype=torch.int32, device=self.device)
            r.counters = torch.zeros(4, dtype=torch.int32, device=self.device)
            r.color = torch.empty((3, req.H, req.W), dtype=torch.float32, device=self.device)
            r.depth = torch.empty((1, req.H, req.W), dtype=torch.float32, device=self.device)
            _lib.check(lib.gs_preprocess(ctypes.byref(r.frame), ctypes.byref(r.gauss), _raster._ptr(r.radii),
                                         _raster._ptr(r.geom), r.geom.numel(), _raster._ptr(r.counters),
                                         ctypes.c_void_p(r.stream.cuda_stream)), "gs_preprocess")
            r.host = torch.empty(4, dtype=torch.int32, pin_memory=True)
            r.host.copy_(r.counters, non_blocking=True)
            r.event = torch.cuda.Event()
            r.event.record(r.stream)
        return r

    def _finish_render_launch(self, r: _Render, num_rendered: int) -> None:
        lib = _lib.load()
        sc = self.scenes[r.req.scene]
        cap = self._round_cap(num_rendered)
        self.seen_counts[(r.req.scene, r.req.W, r.req.H)] = num_rendered
        LB = _raster._binning_layout(cap, sc.P, r.req.W, r.req.H, self.binning)
        # the reservation now uses the real list length
        new = self._render_bytes(r.req, cap)
        self.used += new - r.req.reserved
        r.req.reserved = new
        self.peak_used = max(self.peak_used, self.used)
        with torch.cuda.stream(r.stream):
            r.binning = torch.empty(max(LB.total, 256), dtype=torch.uint8, device=self.device)
            _lib.check(lib.gs_render(ctypes.byref(r.frame), sc.P, _raster._ptr(r.geom), _raster._ptr(r.counters),
                                     cap, _raster._ptr(r.binning), r.binning.numel(), _raster._ptr(r.image),
                                     r.image.numel(), _raster._ptr(r.color), _raster._ptr(r.depth), None,
                                     ctypes.c_void_p(r.stream.cuda_stream)), "gs_render")
            r.event = torch.cuda.Event()
            r.event.record(r.stream)
        r.launched = True

    def _start_solve(self, req: Request, min_capacity: int = 0) -> _Solve:
        # the frame is built on a stream of its own: its target render and the
        # start-capacity forward read list lengths back, and on a shared
        # request stream they would wait behind other requests' work.  Its
        # tensors are released only after the solve's last event, so the
        # solve stream may use them without further bookkeeping.
        ing = self.ingest_stream
        ing.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(ing):
            # one Gaussian per pixel: the frame is ingested on the device from
            # its depth map and image; a subsampled or super-sampled frame
            # (req.P given) takes the host lift
            if req.P is None:
                inst = build_instance_device(req.seed, W=req.W, H=req.H, device=self.device)
            else:
                inst = build_instance(req.seed, W=req.W, H=req.H, P=req.P, device=self.device)
            adam = SE3Adam(1, self.device, lr=self.lr, eps=EPS)
            cap = max(start_capacities([inst], adam.pose, self.binning, self.headroom)[0], int(min_capacity))
            slot = make_slot(inst, cap, self.binning, adam.pose.data_ptr(), adam.grad.data_ptr())
            out3 = torch.zeros(3, dtype=torch.float32, device=self.device)
        st = self._stream()
        st.wait_stream(ing)
        actual = int(slot_bytes(inst.P, inst.W, inst.H, cap, self.binning) + 4 * 14 * inst.P
                     + 4 * 3 * inst.W * inst.H)
        self.used += actual - req.reserved
        req.reserved = actual
        self.peak_used = max(self.peak_used, self.used)
        return _Solve(req=req, stream=st, inst=inst, slot=slot, adam=adam, out3=out3)

    def _launch_solve_iters(self, s: _Solve, n: int) -> None:
        with torch.cuda.stream(s.stream):
            sp = ctypes.c_void_p(s.stream.cuda_stream)
            for _ in range(n):
                launch_slot(s.slot, s.out3.data_ptr(), self.one.data_ptr(), 0.2, sp)
                s.adam.step()
            ev = torch.cuda.Event()
            ev.record(s.stream)
            s.chunks.append(ev)
            if s.iters_done + n >= self.solve_iters:
                # the sticky status word, for the capacity check at the end
                s.host = torch.empty(4, dtype=torch.int32, pin_memory=True)
                s.host.copy_(s.slot.counters, non_blocking=True)
                s.final = torch.cuda.Event()
                s.final.record(s.stream)
        s.iters_done += n

    # -- the loop ----------------------------------------------------------------
    def run(self, requests: Sequence[Request], realtime: bool = True) -> Dict[str, float]:
        """Serves ``requests`` (arrival times relative to the start).  With
        ``realtime`` the clock is the wall clock; otherwise every request is
        treated as arrived at the start (a closed-loop capacity run)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue: deque = deque()
        renders: List[_Render] = []
        solves: List[_Solve] = []
        done: List[Request] = []
        rejected: List[Request] = []
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def retire(req: Request) -> None:
            req.done = now()
            self.used -= req.reserved
            self.max_request_bytes = max(self.max_request_bytes, int(req.reserved))
            done.append(req)

        while pending or queue or renders or solves:
            progressed = False
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                queue.append(pending.popleft())
            # admission, in arrival order
            while queue:
                req = queue[0]
                need = self._render_bytes(req) if req.kind == "render" else self._solve_bytes(req)
                if req.kind == "solve" and len(solves) >= self.max_slots:
                    break
                if self.used + need > self.budget:
                    if not solves and not renders and self.used - self.resident < 1:
                        # nothing in flight and it still does not fit: it never will
                        queue.popleft()
                        req.error = f"needs {need} bytes beyond the resident scenes; budget has " \
                                    f"{self.budget - self.resident:.0f}"
                        req.done = now()
                        rejected.append(req)
                        continue
                    req.waited_for_memory = True
                    break
                queue.popleft()
                req.reserved = need
                self.used += need
                self.peak_used = max(self.peak_used, self.used)
                req.admitted = now()
                if req.kind == "render":
                    renders.append(self._start_render(req))
                else:
                    solves.append(self._start_solve(req))
                progressed = True
            # renders: binning + blend once the list length has landed and its
            # workspace fits the budget, retire when drawn.  A render whose
            # real workspace does not fit waits (it counts as having waited
            # for memory) until others release theirs; the oldest waiting one
            # goes ahead when nothing else could release any.
            still_r = []
            waiting = [r for r in renders if not r.launched and r.event.query()]
            others_busy = bool(solves) or any(r.launched for r in renders)
            for r in renders:
                if r.event.query():
                    if not r.launched:
                        n = int(r.host[0]) & 0xFFFFFFFF
                        extra = self._render_bytes(r.req, self._round_cap(n)) - r.req.reserved
                        # (an event may complete between the two queries: such a
                        # render counts as the oldest waiting one when none was)
                        oldest = waiting[0] if waiting else r
                        if self.used + extra > self.budget and (others_busy or r is not oldest):
                            r.req.waited_for_memory = True
                            still_r.append(r)
                            continue
                        progressed = True
                        self._finish_render_launch(r, n)
                        others_busy = True
                    else:
                        progressed = True
                        r.req.result = r.color
                        retire(r.req)
                        continue
                still_r.append(r)
            renders = still_r
            # solves: keep two chunks queued; retire after the last one
            still_s = []
            for s in solves:
                while s.chunks and s.chunks[0].query():
                    s.chunks.popleft()
                    progressed = True
                if s.iters_done < self.solve_iters and len(s.chunks) < 2:
                    self._launch_solve_iters(s, min(self.iters_per_tick, self.solve_iters - s.iters_done))
                    progressed = True
                if s.iters_done >= self.solve_iters and s.final is not None and s.final.query():
                    progressed = True
                    if int(s.host[3]) & _lib.GS_STATUS_CAPACITY:
                        # binning capacity exceeded during the solve: grow and start over
                        self.used -= s.req.reserved
                        s.req.reserved = 0
                        still_s.append(self._start_solve(s.req, min_capacity=int(s.slot.capacity * 1.5)))
                        continue
                    s.req.result = s.adam.pose[0].clone()
                    s.req.loss = float(s.out3[0])
                    retire(s.req)
                    continue
                still_s.append(s)
            solves = still_s
            self.ticks += 1
            if not progressed:
                # nothing landed and nothing new arrived: wait briefly
                nxt = pending[0].arrival - now() if (realtime and pending and not queue) else 0.0
                time.sleep(min(max(nxt, 0.0), 0.002) if nxt > 0 else 0.0001)
        torch.cuda.synchronize(self.device)
        wall = now()
        out = summarize(done, wall)
        out["rejected"] = len(rejected)
        return out


def summarize(done: Sequence[Request], wall: float) -> Dict[str, float]:
    def pct(xs, q):
        if not xs:
            return float("nan")
        xs = sorted(xs)
        k = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
        return xs[k]

    lat = [r.latency for r in done]
    out = dict(requests=len(done), wall_s=wall, throughput_rps=len(done) / wall if wall > 0 else float("nan"),
               p50_s=pct(lat, 50), p99_s=pct(lat, 99),
               waited_for_memory=sum(r.waited_for_memory for r in done) / max(1, len(done)))
    for kind in ("render", "solve"):
        lk = [r.latency for r in done if r.kind == kind]
        out[f"{kind}_n"] = len(lk)
        out[f"{kind}_p50_s"] = pct(lk, 50)
        out[f"{kind}_p99_s"] = pct(lk, 99)
    return out


def make_requests(n: int, rate: float, n_scenes: int, seed: int = 0, render_frac: float = 0.7,
                  render_res=((480, 270), (960, 540), (1920, 1080)), solve_res=(480, 270),
                  solve_P: Optional[int] = None) -> List[Request]:
    """Poisson arrivals at ``rate`` per second: ``render_frac`` renders of a
    random scene at a random resolution, the rest pose solves of new frames."""
    rng = random.Random(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.expovariate(rate) if rate > 0 else 0.0
        if rng.random() < render_frac:
            W, H = render_res[rng.randrange(len(render_res))]
            reqs.append(Request(rid=i, kind="render", arrival=t, scene=rng.randrange(n_scenes), W=W, H=H,
                                seed=rng.randrange(1 << 30)))
        else:
            reqs.append(Request(rid=i, kind="solve", arrival=t, W=solve_res[0], H=solve_res[1],
                                seed=rng.randrange(1 << 30), P=solve_P))
    return reqs
