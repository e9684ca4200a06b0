"""Pose solves at the workloads' own sizes.

* One frame pair, 480x270, one Gaussian per pixel (129,600): the GPU pose
  gradient against the oracle's, in fp32 and in fp64 (the fp64 backward is
  pinned by central differences in tests/test_oracle_gradients.py), at the
  start pose and at a pose part-way to the hidden one.
* A 10-iteration solve on the plain native path against an oracle-driven loop
  on the CPU: oracle forward and backward, the reference's loss formula in
  torch float64, torch's SE(3) map and torch.optim.Adam on the 6-D tangent
  (trainer/cf3dgs_trainer.py:261-285, scene/gaussian_model_cf.py:303-340).
  The two end within 1e-4 rad / 1e-4 of each other.
* Batched solves of 8 instances spanning every resolution of the batched
  workload (320x180 ... 960x540) and Gaussian counts up to 200k, including a
  count above the pixel count (2x super-grid lift): the graph-captured
  batched solver ends bit for bit on the plain one-solve-per-launch path's
  poses after the full 300 iterations.
"""
import numpy as np
import pytest
import torch

from gpu_helpers import DEV, assert_close_grad, np_
from gs_hip import _lib, raster
from gs_hip.batch import BatchedPoseSolver
from gs_hip.loss import fused_photometric_value_and_grad, photometric_loss
from gs_hip.pose import EPS, LR, build_instance, pose_errors, se3_exp, solve_plain_native
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    print("C3 instance: 480x270, one Gaussian per pixel", flush=True)
    inst = build_instance(3, W=480, H=270)
    assert inst.P == 129_600
    return inst


def _oracle_scene(inst):
    return dict(means=np_(inst.means), shs=np_(inst.shs), opacities=np_(inst.opacities),
                scales=np_(inst.scales), rotations=np_(inst.rotations))


def _oracle_cam(inst, pose7):
    """The instance's raster settings (view = identity, active SH degree 0)
    as the oracle's camera dict, with the pose applied to the means."""
    s = inst.settings
    return dict(W=inst.W, H=inst.H, tanfovx=s.tanfovx, tanfovy=s.tanfovy,
                view=np_(s.viewmatrix).reshape(-1), proj=np_(s.projmatrix).reshape(-1),
                campos=np_(s.campos).reshape(-1), bg=np_(s.bg).astype(np.float32).reshape(-1),
                sh_degree=int(s.sh_degree), scale_modifier=float(s.scale_modifier),
                pose7=np_(pose7).astype(np.float32).reshape(-1))


@pytest.mark.parametrize("xi_scale", [0.0, 0.03])
def test_c3_pose_gradient_matches_oracle(c3, xi_scale):
    print("C3 pose gradient, xi scale", xi_scale, flush=True)
    inst = c3
    xi = torch.randn(6, generator=torch.Generator().manual_seed(5), dtype=torch.float64) * xi_scale
    pose = se3_exp(xi).float().to(DEV)
    color, radii, _, _, st = raster.forward(inst.settings, inst.means, inst.shs, None, inst.opacities, inst.scales,
                                            inst.rotations, None, pose=pose, want_alpha=False)
    _, dcolor = fused_photometric_value_and_grad(color, inst.target, 0.2, clamp01=True)
    g = raster.backward(inst.settings, st, inst.means, inst.shs, None, inst.opacities, inst.scales,
                        inst.rotations, None, pose, dcolor, want={"pose"}, deterministic=True)
    torch.cuda.synchronize()
    scene, cam = _oracle_scene(inst), _oracle_cam(inst, pose)
    o32 = O.forward(scene, cam)
    assert np.array_equal(np_(radii), o32["radii"]) and st.num_rendered == o32["num_rendered"]
    assert float(np.abs(np_(color) - o32["color"]).max()) <= 1e-5
    dc = np_(dcolor)
    g32 = O.backward(scene, cam, o32, dc)["pose7"]
    o64 = O.forward(scene, cam, np.float64)
    g64 = O.backward(scene, cam, o64, dc.astype(np.float64), dtype=np.float64)["pose7"]
    got = np_(g["pose"])
    print("gpu", got, "\noracle fp32", g32, "\noracle fp64", g64, flush=True)
    assert_close_grad("pose vs fp32 oracle", got, g32, rtol=2e-3, atol_frac=2e-4)
    assert_close_grad("pose vs fp64 oracle", got, g64, rtol=5e-3, atol_frac=1e-3)


def test_c3_solve_matches_oracle_driven_loop(c3):
    iters = 10
    print("C3 solve: plain native vs an oracle-driven loop,", iters, "iterations", flush=True)
    inst = c3
    gpu_pose = solve_plain_native(inst, iters=iters)["pose"].cpu()
    scene = _oracle_scene(inst)
    target = inst.target.cpu().double()
    xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([xi], lr=LR, eps=EPS)
    for it in range(iters):
        with torch.no_grad():
            pose = se3_exp(xi).float()
        cam = _oracle_cam(inst, pose)
        o = O.forward(scene, cam)
        img = torch.from_numpy(o["color"]).double().requires_grad_(True)
        photometric_loss(img.clamp(0.0, 1.0), target, 0.2).backward()
        g7 = O.backward(scene, cam, o, img.grad.float().numpy())["pose7"]
        opt.zero_grad()
        (se3_exp(xi) * torch.from_numpy(np.asarray(g7, np.float64))).sum().backward()
        opt.step()
    with torch.no_grad():
        ref_pose = se3_exp(xi).float()
    rot, trans = pose_errors(gpu_pose, ref_pose)
    moved = pose_errors(ref_pose, torch.tensor([0.0, 0, 0, 0, 0, 0, 1]))
    print("gpu vs oracle loop", rot, trans, "moved", moved, flush=True)
    assert max(moved) > 1e-3  # the loop did move the pose
    assert rot <= 1e-4 and trans <= 1e-4


C4_SUBSET = [(320, 180, 20_000), (480, 270, 129_600), (640, 360, 200_000), (960, 540, 200_000),
             (320, 180, 80_000), (480, 270, 60_000), (640, 360, 110_000), (960, 540, 45_000)]


def test_c4_subset_batched_equals_plain_at_full_size():
    print("C4 subset: 8 instances, 300 iterations, batched vs plain", flush=True)
    insts = [build_instance(1000 + k, W=W, H=H, P=P) for k, (W, H, P) in enumerate(C4_SUBSET)]
    assert insts[4].supergrid and max(i.P for i in insts) == 200_000
    plain = [solve_plain_native(i, iters=300)["pose"] for i in insts]
    # the schedule the autotuner picks for the batched workload: per-frame
    # launches on 8 streams
    solver = BatchedPoseSolver(insts, iters=300, streams=8, binning=_lib.BINNING_DEPTH_FIRST, iters_per_graph=10,
                               multi=False)
    poses, losses = solver.run()
    assert torch.isfinite(losses).all()
    for k, inst in enumerate(insts):
        print(k, C4_SUBSET[k], "vs plain", pose_errors(poses[k], plain[k]), "vs gt",
              pose_errors(poses[k], inst.gt_pose), flush=True)
        assert torch.equal(poses[k], plain[k]), k
