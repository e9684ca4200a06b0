"""Serving loop on the GPU: results in a mix equal results alone; memory
admission; latency bookkeeping."""
import copy

import pytest
import torch

from gs_hip.serve import Request, Server, make_requests, make_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _scenes():
    return [make_scene(k, P, DEV) for k, P in enumerate((20000, 50000, 80000))]


def _mix():
    reqs = []
    res = [(160, 90), (320, 180), (240, 135)]
    for i in range(8):
        W, H = res[i % 3]
        reqs.append(Request(rid=i, kind="render", arrival=0.0, scene=i % 3, W=W, H=H, seed=100 + i))
    for i in range(3):
        reqs.append(Request(rid=8 + i, kind="solve", arrival=0.0, W=128, H=72, seed=200 + i))
    return reqs


def test_mix_equals_alone_and_memory_admission():
    print("serve: mix vs alone", flush=True)
    scenes = _scenes()
    kw = dict(solve_iters=30, iters_per_tick=5, streams=3)
    probe = Server(scenes, **kw)
    need = [probe._render_bytes(r) if r.kind == "render" else probe._solve_bytes(r) for r in _mix()]
    # room for the resident scenes plus about two of the largest requests: later ones wait for memory
    budget = probe.resident + 2.5 * max(need)
    mix = _mix()
    srv = Server(scenes, budget_bytes=budget, **kw)
    stats = srv.run(mix, realtime=False)
    print(stats, flush=True)
    assert stats["requests"] == len(mix) and stats["rejected"] == 0
    assert stats["waited_for_memory"] > 0
    assert srv.peak_used <= budget * 1.05
    assert abs(srv.used - srv.resident) < 1  # every reservation released
    for r in mix:
        alone = copy.copy(r)
        alone.result = None
        alone.done = None
        alone.waited_for_memory = False
        Server(scenes, **kw).run([alone], realtime=False)
        assert torch.equal(alone.result, r.result), (r.kind, r.rid)
        assert r.latency >= 0


def test_poisson_arrivals_realtime():
    print("serve: poisson", flush=True)
    scenes = _scenes()
    reqs = make_requests(12, rate=40.0, n_scenes=len(scenes), seed=3,
                         render_res=((160, 90), (320, 180)), solve_res=(128, 72))
    srv = Server(scenes, solve_iters=20, iters_per_tick=5, streams=2)
    stats = srv.run(reqs, realtime=True)
    print(stats, flush=True)
    assert stats["requests"] == 12
    assert stats["p99_s"] >= stats["p50_s"] > 0
    for r in reqs:
        assert r.done >= r.admitted >= r.arrival - 1e-6
        assert r.result is not None and torch.isfinite(r.result).all()


def test_request_larger_than_budget_is_rejected():
    scenes = _scenes()
    probe = Server(scenes)
    big = Request(rid=0, kind="render", arrival=0.0, scene=2, W=320, H=180, seed=1)
    small = Request(rid=1, kind="render", arrival=0.0, scene=0, W=160, H=90, seed=2)
    budget = probe.resident + 0.5 * probe._render_bytes(big) + probe._render_bytes(small)
    budget = max(budget, probe.resident + probe._render_bytes(small) * 1.01)
    stats = Server(scenes, budget_bytes=budget).run([big, small], realtime=False)
    assert stats["rejected"] == 1 and big.error and big.result is None
    assert small.result is not None


def test_mix_equals_alone_with_a_large_scene_at_1080p():
    """The serving workload's large end: a 1.2M-Gaussian scene rendered at
    1920x1080 and 960x540 beside a 200k scene, with a 480x270 pose solve of
    200 iterations in the same ticks.  Every result equals its run alone bit
    for bit."""
    print("serve: 1.2M scene at 1920x1080 in a mix", flush=True)
    scenes = [make_scene(0, 1_200_000, DEV), make_scene(1, 200_000, DEV)]
    reqs = [Request(rid=0, kind="render", arrival=0.0, scene=0, W=1920, H=1080, seed=11),
            Request(rid=1, kind="solve", arrival=0.0, W=480, H=270, seed=12),
            Request(rid=2, kind="render", arrival=0.0, scene=1, W=1920, H=1080, seed=13),
            Request(rid=3, kind="render", arrival=0.0, scene=0, W=960, H=540, seed=14),
            Request(rid=4, kind="render", arrival=0.0, scene=1, W=480, H=270, seed=15)]
    kw = dict(solve_iters=200, iters_per_tick=10, streams=4)
    stats = Server(scenes, **kw).run(reqs, realtime=False)
    print(stats, flush=True)
    assert stats["requests"] == len(reqs) and stats["rejected"] == 0
    for r in reqs:
        alone = copy.copy(r)
        alone.result = None
        alone.done = None
        alone.waited_for_memory = False
        Server(scenes, **kw).run([alone], realtime=False)
        assert torch.equal(alone.result, r.result), (r.kind, r.rid)
