"""Pose solve on the GPU: SE(3) kernels, the plain solver, the graph-captured
batched solver.

* gs_se3_retract / gs_se3_grad against the torch float64 SE(3) map (itself
  pinned to the matrix exponential on the CPU) and its autograd derivative;
* gs_se3_adam against torch.optim.Adam on the tangent with that gradient;
* the tangent gradient of the whole render -> loss chain through autograd
  against the kernel Jacobian applied to the pose gradient, and its
  direction against central differences of the rendered loss;
* plain solves converge to the hidden pose; the batched solver (graph,
  several streams, both binning schedules) ends bit for bit on the pose of
  the plain one-solve-per-launch path (fixed-order reductions), so well
  inside 1e-4 rad / 1e-4 translation;
* every graph replay agrees bit for bit with eager launches of the same
  iterations;
* a binning-capacity overflow (forced by a small start capacity) is caught by
  the sticky device flag and re-solved with a larger capacity.
"""
import pytest
import torch

from gs_hip import _lib
from gs_hip.batch import BatchedPoseSolver
from gs_hip.loss import fused_photometric_loss
from gs_hip.pose import (SE3Adam, build_instance, pose_errors, se3_exp, se3_grad, se3_retract, solve_plain,
                         solve_plain_native)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _rand_xi(n, scale, seed):
    g = torch.Generator().manual_seed(seed)
    xi = torch.randn((n, 6), generator=g, dtype=torch.float64) * scale
    xi[0] = 0.0
    return xi


@pytest.mark.parametrize("with_x0", [False, True])
def test_se3_retract_and_grad_match_torch(with_x0):
    print("se3 kernels, X0:", with_x0, flush=True)
    xi = torch.cat([_rand_xi(8, 0.3, 1), _rand_xi(4, 1e-4, 2), _rand_xi(4, 1.5, 3)])
    n = xi.shape[0]
    x0 = se3_exp(_rand_xi(n, 0.6, 4)) if with_x0 else None
    got = se3_retract(xi.float().to(DEV), x0.float().to(DEV) if x0 is not None else None).cpu().double()
    ref = se3_exp(xi.float().double(), x0.float().double() if x0 is not None else None)
    torch.testing.assert_close(got, ref, rtol=0, atol=2e-6)
    gp = torch.randn((n, 7), generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    gx = se3_grad(xi.float().to(DEV), gp.float().to(DEV), x0.float().to(DEV) if x0 is not None else None)
    xr = xi.float().double().requires_grad_(True)
    (se3_exp(xr, x0.float().double() if x0 is not None else None) * gp.float().double()).sum().backward()
    torch.testing.assert_close(gx.cpu().double(), xr.grad, rtol=1e-5, atol=1e-6)


def test_se3_adam_matches_torch_adam():
    n = 5
    g = torch.Generator().manual_seed(7)
    x0 = se3_exp(_rand_xi(n, 0.5, 8)).float()
    opt = SE3Adam(n, DEV, lr=1e-3, X0=x0.to(DEV))
    xi_ref = torch.zeros((n, 6), dtype=torch.float32, requires_grad=True)
    ref = torch.optim.Adam([xi_ref], lr=1e-3, eps=1e-15)
    torch.testing.assert_close(opt.pose.cpu(), x0, rtol=0, atol=1e-6)
    for it in range(6):
        gp = torch.randn((n, 7), generator=g)
        opt.grad.copy_(gp.to(DEV))
        opt.step()
        ref.zero_grad()
        xd = xi_ref.detach().double().requires_grad_(True)
        (se3_exp(xd, x0.double()) * gp.double()).sum().backward()
        xi_ref.grad = xd.grad.float()
        ref.step()
        torch.testing.assert_close(opt.xi.cpu(), xi_ref.detach(), rtol=1e-5, atol=1e-7)
        with torch.no_grad():
            pref = se3_exp(xi_ref.double(), x0.double()).float()
        torch.testing.assert_close(opt.pose.cpu(), pref, rtol=0, atol=2e-6)
    assert int(opt.step_count.item()) == 6


def _loss_at(inst, xi):
    from diff_gaussian_rasterization import GaussianRasterizer
    r = GaussianRasterizer(raster_settings=inst.settings)
    pose = se3_exp(xi.to(DEV).float())
    color, _, _ = r(means3D=inst.means, means2D=torch.zeros_like(inst.means), opacities=inst.opacities,
                    shs=inst.shs, scales=inst.scales, rotations=inst.rotations, pose=pose)
    loss, _ = fused_photometric_loss(color, inst.target, 0.2, clamp01=True)
    return loss


def test_pose_tangent_gradient_through_autograd():
    """xi -> se3_exp (torch) -> drop-in rasteriser -> fused loss: the tangent
    gradient autograd delivers equals gs_se3_grad applied to the pose
    gradient of the same render (the pose gradient itself is pinned against
    the oracle's FP64 finite differences elsewhere)."""
    print("pose tangent gradient via autograd", flush=True)
    inst = build_instance(11, W=160, H=90)
    xi = torch.tensor([0.01, -0.02, 0.015, 0.02, -0.01, 0.005], dtype=torch.float32)
    xr = xi.clone().to(DEV).requires_grad_(True)
    _loss_at(inst, xr).backward()
    from diff_gaussian_rasterization import GaussianRasterizer
    r = GaussianRasterizer(raster_settings=inst.settings)
    pose = se3_exp(xi.to(DEV)).detach().requires_grad_(True)
    color, _, _ = r(means3D=inst.means, means2D=torch.zeros_like(inst.means), opacities=inst.opacities,
                    shs=inst.shs, scales=inst.scales, rotations=inst.rotations, pose=pose)
    fused_photometric_loss(color, inst.target, 0.2, clamp01=True)[0].backward()
    via_kernel = se3_grad(xi.to(DEV), pose.grad)
    torch.testing.assert_close(xr.grad.reshape(1, 6), via_kernel, rtol=2e-3, atol=2e-3 * float(via_kernel.abs().max()))


def test_lietorch_compat_path_matches_fused_pose():
    """The reference's unchanged code shape -- a LieGroupParameter and
    ``P.retr().act(xyz)`` in torch feeding the drop-in rasteriser
    (scene/gaussian_model_cf.py:123-133) -- renders the image of the fused-pose
    path (pose applied inside preprocess) and delivers the same tangent
    gradient."""
    from diff_gaussian_rasterization import GaussianRasterizer
    from lietorch import SE3, LieGroupParameter
    print("lietorch compat vs fused pose", flush=True)
    inst = build_instance(11, W=160, H=90)
    xi = torch.tensor([0.01, -0.02, 0.015, 0.02, -0.01, 0.005], dtype=torch.float32)
    P = LieGroupParameter(SE3(torch.tensor([0.0, 0, 0, 0, 0, 0, 1.0], device=DEV)[None]))
    with torch.no_grad():
        P.copy_(xi[None].to(DEV))
    means = P.retr().act(inst.means.clone())
    r = GaussianRasterizer(raster_settings=inst.settings)
    color_a, _, _ = r(means3D=means, means2D=torch.zeros_like(means), opacities=inst.opacities, shs=inst.shs,
                      scales=inst.scales, rotations=inst.rotations)
    fused_photometric_loss(color_a, inst.target, 0.2, clamp01=True)[0].backward()
    xr = xi.clone().to(DEV).requires_grad_(True)
    color_b, _, _ = r(means3D=inst.means, means2D=torch.zeros_like(inst.means), opacities=inst.opacities,
                      shs=inst.shs, scales=inst.scales, rotations=inst.rotations, pose=se3_exp(xr))
    fused_photometric_loss(color_b, inst.target, 0.2, clamp01=True)[0].backward()
    torch.testing.assert_close(color_a, color_b, rtol=0, atol=1e-5)
    ga, gb = P.grad.reshape(6), xr.grad
    torch.testing.assert_close(ga, gb, rtol=2e-3, atol=2e-3 * float(gb.abs().max()))


def test_pose_tangent_gradient_direction_matches_fd():
    """Central differences of the rendered loss in the tangent (h = 1e-3).
    The renderer truncates every splat (alpha < 1/255, 3-sigma rect), and the
    analytic gradient -- the reference's as well as this one -- is the
    derivative of the untruncated interior only: the truncation boundary
    adds a first-order term to any finite difference of the image, so the
    magnitudes differ by design.  The direction must agree."""
    inst = build_instance(11, W=160, H=90)
    xi = torch.tensor([0.01, -0.02, 0.015, 0.02, -0.01, 0.005], dtype=torch.float32)
    xr = xi.clone().to(DEV).requires_grad_(True)
    _loss_at(inst, xr).backward()
    grad = xr.grad.cpu().double()
    h = 1e-3
    fd = torch.zeros(6, dtype=torch.float64)
    for k in range(6):
        e = torch.zeros(6)
        e[k] = h
        with torch.no_grad():
            fd[k] = (float(_loss_at(inst, xi + e)) - float(_loss_at(inst, xi - e))) / (2 * h)
    cos = float(grad @ fd) / float(grad.norm() * fd.norm())
    print("grad", grad.tolist(), "fd", fd.tolist(), "cos", cos, flush=True)
    assert cos > 0.98 and (torch.sign(grad) == torch.sign(fd)).all()


def test_plain_solve_converges():
    print("plain solve", flush=True)
    inst = build_instance(21, W=160, H=90)
    out = solve_plain(inst, iters=300, fused_loss=True)
    rot, trans = pose_errors(out["pose"], inst.gt_pose)
    rot0, trans0 = pose_errors(torch.tensor([0.0, 0, 0, 0, 0, 0, 1]), inst.gt_pose)
    print("start", rot0, trans0, "end", rot, trans, "loss", out["loss"], flush=True)
    assert rot < 0.1 * rot0 and trans < 0.1 * trans0


@pytest.mark.parametrize("binning,groups,shared", [(_lib.BINNING_DEPTH_FIRST, 1, False),
                                                   (_lib.BINNING_DEPTH_FIRST, 2, False),
                                                   (_lib.BINNING_SORT64, 0, False),
                                                   (_lib.BINNING_DEPTH_FIRST, 0, True),
                                                   (_lib.BINNING_DEPTH_FIRST, 2, True)])
def test_batched_equals_plain_native(binning, groups, shared):
    """Batched (graph, 2 streams; blend and render backward in one launch
    over all frames or over each of two instance groups, or per frame; the
    depth sorts per frame or as one segmented sort per stream or group) vs
    the plain one-solve-per-launch path on the same kernels: every reduction
    is in a fixed order, so the final poses agree bit for bit (well inside
    1e-4 rad / 1e-4)."""
    print("batched vs plain native, binning", binning, "groups", groups, "shared sort", shared, flush=True)
    insts = [build_instance(31 + k, W=W, H=H, P=P) for k, (W, H, P) in
             enumerate([(160, 90, 12000), (192, 108, 30000), (128, 72, 9216)])]
    plain = [solve_plain_native(i, iters=300)["pose"] for i in insts]
    solver = BatchedPoseSolver(insts, iters=300, streams=2, binning=binning, iters_per_graph=10, multi=groups > 0,
                               groups=max(groups, 1), shared_sort=shared)
    poses, losses = solver.run()
    assert torch.isfinite(losses).all()
    for k, inst in enumerate(insts):
        rot, trans = pose_errors(poses[k], plain[k])
        print(k, "batched vs plain", rot, trans, "vs gt", pose_errors(poses[k], inst.gt_pose), flush=True)
        assert torch.equal(poses[k], plain[k])
        assert rot <= 1e-4 and trans <= 1e-4


def test_native_and_torch_plain_reach_the_same_pose():
    """The native plain path and the torch-op path (torch SE(3), torch loss,
    torch Adam) end within the optimiser's own jitter of each other and of
    the hidden pose."""
    inst = build_instance(61, W=160, H=90)
    a = solve_plain_native(inst, iters=300)["pose"]
    b = solve_plain(inst, iters=300)["pose"]
    ra, ta = pose_errors(a, inst.gt_pose)
    rb, tb = pose_errors(b, inst.gt_pose)
    rab, tab = pose_errors(a, b)
    print("native vs gt", ra, ta, "torch vs gt", rb, tb, "native vs torch", rab, tab, flush=True)
    assert max(ra, ta, rb, tb) < 2e-3 and max(rab, tab) < 2e-3


_SIDE_TO_SIDE = pytest.mark.skip(
    reason="two streams per instance group made one side stream wait on another inside the graph capture; "
           "hipStreamEndCapture segfaulted on it twice on this ROCm build, and the solver no longer builds that "
           "topology (several groups take one stream each)")


@pytest.mark.parametrize("groups,streams,shared", [(0, 3, False), (1, 3, False), (2, 2, False), (0, 3, True),
                                                   (1, 3, True),
                                                   pytest.param(2, 3, False, marks=_SIDE_TO_SIDE)])
def test_graph_replays_match_eager(groups, streams, shared):
    """Every graph replay agrees bit for bit with eager launches of the same
    iterations: per-frame launches on streams (the instances dealt to the
    streams by cost), or one-launch blend and backward over all instances or
    over each of two instance groups, each group on a stream of its own; the
    depth sorts per frame or as one segmented sort per stream or group."""
    print("graph replays vs eager, groups", groups, "streams", streams, "shared sort", shared, flush=True)
    insts = [build_instance(41 + k, W=128, H=72) for k in range(3)] + [build_instance(44, W=160, H=90)]
    ipg = 4
    kw = dict(multi=groups > 0, groups=max(groups, 1), shared_sort=shared)
    eager = BatchedPoseSolver(insts, iters=ipg, streams=streams, use_graph=False, **kw)
    graph = BatchedPoseSolver(insts, iters=ipg, streams=streams, iters_per_graph=ipg, **kw)
    graph._reset()
    graph._capture()
    graph._reset()
    eager._reset()
    for rep in range(4):
        graph.graph.replay()
        eager._launch_iterations(ipg)
        torch.cuda.synchronize()
        # host work and eager launches between replays
        assert int(graph.adam.step_count.item()) == ipg * (rep + 1)
        assert torch.equal(graph.adam.pose, eager.adam.pose)
        assert torch.equal(graph.out3, eager.out3)


@pytest.mark.parametrize("groups,shared", [(1, False), (2, True)])
def test_capacity_overflow_regrows(groups, shared):
    """A binning capacity too small for some instance raises the sticky flag;
    the solver grows those instances and solves again, with the same result.
    With shared stages the regrowth also rebuilds the shared sorts' and the
    loss's sets over the new buffers."""
    print("capacity overflow regrow, groups", groups, "shared", shared, flush=True)
    insts = [build_instance(51 + k, W=128, H=72) for k in range(2)]
    kw = dict(streams=groups, iters_per_graph=5, groups=groups, shared_sort=shared)
    ref_solver = BatchedPoseSolver(insts, iters=20, **kw)
    ref_poses, _ = ref_solver.run()
    assert ref_solver.regrows == 0
    small = BatchedPoseSolver(insts, iters=20, headroom=0.3, **kw)
    poses, _ = small.run()
    assert small.regrows >= 1
    assert torch.equal(poses, ref_poses)


@pytest.mark.parametrize("W,H,P", [(160, 90, 14400), (480, 270, 129_600)])
def test_pose_only_backward_matches_the_full_backward_bitwise(W, H, P):
    """The pose-only backward (only the pose gradient wanted, fixed-order,
    colour loss, SH degree 0) sums five terms per pair instead of ten; its
    pose gradient has the same bits as the full backward's, and the library
    says it takes that path."""
    from gs_hip import raster
    from gs_hip.loss import fused_photometric_value_and_grad
    print("pose-only vs full backward", W, H, P, flush=True)
    inst = build_instance(71, W=W, H=H, P=P)
    pose = se3_exp(torch.tensor([0.01, -0.02, 0.015, 0.02, 0.01, -0.03], dtype=torch.float64)).float().to(DEV)
    color, _, _, _, st = raster.forward(inst.settings, inst.means, inst.shs, None, inst.opacities, inst.scales,
                                        inst.rotations, None, pose=pose, want_alpha=False)
    _, dcolor = fused_photometric_value_and_grad(color, inst.target, 0.2, clamp01=True)
    args = (inst.settings, st, inst.means, inst.shs, None, inst.opacities, inst.scales, inst.rotations, None, pose,
            dcolor)
    full = raster.backward(*args, deterministic=True)
    po = raster.backward(*args, want={"pose"}, deterministic=True)
    torch.cuda.synchronize()
    assert torch.equal(full["pose"], po["pose"]) and torch.count_nonzero(po["pose"]) == 7
    fa = raster.make_frame(inst.settings, DEV, pose, None, True)
    ga = raster.make_gaussians(inst.means, inst.shs, None, inst.opacities, inst.scales, inst.rotations, None, DEV)
    g = _lib.GsGrads()
    g.dL_dpose = po["pose"].data_ptr()
    import ctypes
    assert _lib.load().gs_backward_is_pose_only(ctypes.byref(fa.frame), ctypes.byref(ga.g), ctypes.byref(g), 0) == 1
    assert _lib.load().gs_backward_is_pose_only(ctypes.byref(fa.frame), ctypes.byref(ga.g), ctypes.byref(g), 1) == 0
