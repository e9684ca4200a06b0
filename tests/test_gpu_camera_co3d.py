"""The CF pipeline's camera through the drop-in rasteriser on the GPU.

Camera(..., is_co3d=True) (scene/cameras.py:57-66, trainer/trainer.py:716-719)
builds the view with getWorld2View3 and the projection from the intrinsics,
whose principal point puts non-zero terms in the projection's third column.
Here cx != W/2 and cy != H/2 on a rotated, translated camera: the forward's
integers and image must equal the oracle's, and the gradients that
``GaussianRasterizer`` routes to every input (means2D included) must match the
oracle's analytic backward, which tests/test_oracle_branches.py pins by
central differences for this projection.
"""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from gpu_helpers import DEV, assert_close_grad, np_
from gs_hip import scenes
from gs_hip.camera import make_camera_co3d, oracle_camera, raster_settings
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deg", [0, 3])
def test_co3d_camera_forward_backward_through_the_drop_in(deg):
    from diff_gaussian_rasterization import GaussianRasterizer
    W, H = 240, 160
    print("co3d camera, principal point off centre, degree", deg, flush=True)
    sc = scenes.plain_scene(seed=70 + deg, P=6000, W=W, H=H)
    f = W / (2.0 * np.tan(np.radians(30.0)))
    K = np.array([[f, 0.0, 0.43 * W], [0.0, 1.02 * f, 0.58 * H], [0.0, 0.0, 1.0]])
    R = Rotation.from_rotvec([0.03, -0.02, 0.05]).as_matrix()
    cam = make_camera_co3d(R, np.array([0.05, -0.04, 0.1]), K, W, H)
    assert abs(float(cam.projection_matrix[2, 0])) > 0.1 and abs(float(cam.projection_matrix[2, 1])) > 0.1

    dcam = cam.to(DEV)
    rs = raster_settings(dcam, sc["bg"].to(DEV), deg)
    leaves = {k: sc[k].to(DEV).clone().requires_grad_(True)
              for k in ("means", "shs", "opacities", "scales", "rotations")}
    means2D = torch.zeros_like(leaves["means"], requires_grad=True)
    color, radii, depth = GaussianRasterizer(rs)(
        means3D=leaves["means"], means2D=means2D, shs=leaves["shs"], colors_precomp=None,
        opacities=leaves["opacities"], scales=leaves["scales"], rotations=leaves["rotations"], cov3D_precomp=None)
    g = torch.Generator().manual_seed(deg + 5)
    dc = (torch.rand((3, H, W), generator=g) - 0.5) / (H * W)
    dd = (torch.rand((1, H, W), generator=g) - 0.5) / (H * W)
    ((color * dc.to(DEV)).sum() + (depth * dd.to(DEV)).sum()).backward()
    torch.cuda.synchronize()

    ocam = oracle_camera(cam, sc["bg"], deg)
    scene = {"means": sc["means"].numpy(), "shs": sc["shs"].numpy(), "opacities": sc["opacities"].numpy(),
             "scales": sc["scales"].numpy(), "rotations": sc["rotations"].numpy()}
    o = O.forward(scene, ocam)
    assert o["num_rendered"] > 10_000
    assert np.array_equal(np_(radii), o["radii"])
    assert float(np.abs(np_(color) - o["color"]).max()) <= 1e-5
    assert float(np.abs(np_(depth) - o["depth"]).max()) <= 1e-5
    og = O.backward(scene, ocam, o, dc.numpy(), dd.numpy())
    assert_close_grad("means2D", np_(means2D.grad), og["means2D"])
    assert_close_grad("means3D", np_(leaves["means"].grad), og["means3D"])
    assert_close_grad("opacities", np_(leaves["opacities"].grad).reshape(-1), og["opacities"])
    assert_close_grad("scales", np_(leaves["scales"].grad), og["scales"])
    assert_close_grad("rotations", np_(leaves["rotations"].grad), og["rotations"])
    assert_close_grad("shs", np_(leaves["shs"].grad), og["shs"])
