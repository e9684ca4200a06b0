"""Forward parity of the HIP rasteriser against the fp32 CPU oracle.

Bar: integer outputs (radii, tiles touched, duplicated and sorted keys, sorted
Gaussian ids, tile ranges, contributor counts) are bit-exact; the per-Gaussian
screen quantities and the colour / depth / alpha images are bit-exact too
(both sides use the same explicit rounding steps), and in any case within the
stated image tolerance max |error| <= 1e-5.
"""

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, np_, run_gpu, run_oracle, u64
from gs_hip import scenes

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5


def _compare(sc, sh_degree, binning=None, **kw):
    g = run_gpu(sc, sh_degree, binning=binning, **kw)
    _, _, o = run_oracle(sc, sh_degree, **kw)
    v = g["views"]
    radii = np_(g["radii"])
    np.testing.assert_array_equal(radii, o["radii"])
    vis = radii > 0
    np.testing.assert_array_equal(np_(v["tiles"]).view(np.uint32), o["tiles"])
    np.testing.assert_array_equal(np_(v["xy"])[vis], o["xy"][vis])
    np.testing.assert_array_equal(np_(v["depth"])[vis], o["gdepth"][vis])
    np.testing.assert_array_equal(np_(v["conic"])[vis], o["conic_op"][vis, 0:2])
    np.testing.assert_array_equal(np_(v["conic_c"])[vis], o["conic_op"][vis, 2])
    np.testing.assert_array_equal(np_(v["rgb"])[vis], o["rgb"][vis])
    bits = np_(v["clamp_bits"])[vis]
    ob = o["clamped"][vis].astype(np.int32)
    np.testing.assert_array_equal(bits, ob[:, 0] | (ob[:, 1] << 1) | (ob[:, 2] << 2))
    assert g["st"].num_rendered == o["num_rendered"]
    n = o["num_rendered"]
    # (the duplicated, unsorted keys live in the sort's ping-pong buffers and are
    # overwritten by it; their order is pinned on the CPU side, and the stable
    # tie order it implies is visible in the sorted ids below)
    np.testing.assert_array_equal(u64(v["keys"]), o["keys"])
    np.testing.assert_array_equal(np_(v["vals"]).view(np.uint32), o["vals"])
    np.testing.assert_array_equal(np_(v["ranges"]).view(np.uint32), o["ranges"])
    np.testing.assert_array_equal(np_(v["n_contrib"]), o["n_contrib"])
    for name, gt, ot in (("color", g["color"], o["color"]), ("depth", g["depth"], o["depth"]),
                         ("alpha", g["alpha"], o["alpha"])):
        err = float(np.abs(np_(gt) - ot).max())
        assert err <= IMG_TOL, (name, err)
        np.testing.assert_array_equal(np_(gt), ot, err_msg=name)
    np.testing.assert_array_equal(np_(v["final_T"]), o["final_T"])
    return g, o, n


@pytest.mark.parametrize("binning", [0, 1], ids=["depth_first", "sort64"])
def test_plain_scene_forward_bit_exact(binning):
    sc = scenes.plain_scene(seed=0)
    g, o, n = _compare(sc, 3, binning=binning)
    assert n > 10000
    assert float(o["alpha"].mean()) > 0.1


@pytest.mark.parametrize("binning", [0, 1], ids=["depth_first", "sort64"])
def test_training_scene_binning_bit_exact(binning):
    """The 500k-Gaussian 960x540 scene: sorted ids, ranges, contributor counts
    and the images of both binning schedules equal the oracle's."""
    import torch.nn.functional as F
    raw = scenes.training_scene(seed=0)
    sc = dict(means=raw["xyz"], scales=torch.exp(raw["scaling"]), rotations=F.normalize(raw["rotation"]),
              opacities=torch.sigmoid(raw["opacity"]), shs=torch.cat([raw["features_dc"], raw["features_rest"]], 1),
              cam=raw["cam"], bg=raw["bg"])
    g, o, n = _compare(sc, 3, binning=binning)
    assert n > 5_000_000


@pytest.mark.parametrize("W,H", [(320, 180), (600, 300), (1920, 1080)], ids=["240-tiles", "722-tiles", "8160-tiles"])
def test_list_sort_widths_bit_exact(W, H):
    """Depth-first tile lists of 240, 722 and 8160 tiles: one 8-bit pass on
    the tile bits, then two 8-bit passes with a 2- and a 5-bit top digit.
    Sorted ids, ranges and images equal the oracle's in each case."""
    print("list sort widths", W, H, flush=True)
    sc = scenes.plain_scene(seed=21, P=20000, W=W, H=H)
    g, o, n = _compare(sc, 3, binning=0)
    assert n > 20000


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_sh_degrees(deg):
    sc = scenes.plain_scene(seed=3, P=3000)
    _compare(sc, deg)


def test_colors_precomp_and_cov3D_precomp():
    sc = scenes.plain_scene(seed=4, P=3000)
    g = torch.Generator().manual_seed(9)
    colors = torch.rand((3000, 3), generator=g)
    _compare(sc, 3, colors=colors)
    # cov3D from the scales / rotations, as CFGaussianModel.get_covariance builds it
    q = sc["rotations"]
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
                     2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
                     2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], dim=1).view(-1, 3, 3)
    L = R * sc["scales"][:, None, :]
    S = L @ L.transpose(1, 2)
    cov = torch.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], dim=1)
    _compare(sc, 3, cov3D=cov.contiguous())


def test_pose_fused_forward():
    sc = scenes.plain_scene(seed=5, P=4000)
    pose = scenes.random_pose(11, 0.1, 0.1)
    _compare(sc, 3, pose=pose)


def test_ragged_image_and_white_background():
    sc = scenes.plain_scene(seed=6, P=5000, W=250, H=131)
    sc["bg"] = torch.ones(3)
    _compare(sc, 2)


def test_large_gaussians_cover_all_tiles():
    sc = scenes.plain_scene(seed=7, P=64, W=96, H=80)
    sc["scales"] = sc["scales"] * 40.0
    g, o, n = _compare(sc, 1)
    assert n > 64 * 10


def test_empty_and_culled():
    sc = scenes.plain_scene(seed=8, P=200)
    # everything behind the near plane
    sc["means"] = sc["means"].clone()
    sc["means"][:, 2] = -1.0
    g, o, n = _compare(sc, 3)
    assert n == 0 and int((np_(g["radii"]) > 0).sum()) == 0
    assert float(np_(g["alpha"]).max()) == 0.0
    sc0 = scenes.plain_scene(seed=8, P=0)
    g0 = run_gpu(sc0, 3)
    assert g0["radii"].numel() == 0
    np.testing.assert_array_equal(np_(g0["color"]), np.zeros((3, 256, 256), np.float32))


def test_nonfinite_inputs_are_culled():
    sc = scenes.plain_scene(seed=10, P=1000)
    sc["means"] = sc["means"].clone()
    sc["scales"] = sc["scales"].clone()
    sc["means"][0, 0] = float("nan")
    sc["means"][1, 2] = float("inf")
    sc["scales"][2, 1] = float("nan")
    g = run_gpu(sc, 3)
    radii = np_(g["radii"])
    assert radii[0] == 0 and radii[1] == 0 and radii[2] == 0
    assert np.isfinite(np_(g["color"])).all()


def test_prefiltered_flag_raises():
    from gs_hip import camera as gcam, raster
    sc = scenes.plain_scene(seed=12, P=100)
    sc["means"] = sc["means"].clone()
    sc["means"][0, 2] = -1.0
    cam = sc["cam"].to(DEV)
    rs = gcam.raster_settings(cam, sc["bg"].to(DEV), 3, prefiltered=True)
    with pytest.raises(RuntimeError, match="prefiltered"):
        raster.forward(rs, sc["means"].to(DEV), sc["shs"].to(DEV), None, sc["opacities"].to(DEV),
                       sc["scales"].to(DEV), sc["rotations"].to(DEV), None)


@pytest.mark.parametrize("binning", [0, 1])
def test_prepare_resume_equals_one_call(binning):
    """The forward queues the binning work that does not need num_rendered
    under the previous frame's capacity (gs_render_prepare) and resumes after
    the readback (gs_render_resume); a list that outgrew that capacity takes
    the one-call path with a larger workspace.  Both equal a first call, which
    has no previous capacity, bit for bit."""
    from gs_hip import raster
    sc = scenes.plain_scene(seed=11, P=3000, W=200, H=136)
    key = (3000, 200, 136, binning)
    raster._LAST_CAPACITY.pop(key, None)
    first = run_gpu(sc, 3, binning=binning)  # one call: no previous capacity for this size
    assert raster._LAST_CAPACITY[key] == first["st"].capacity
    again = run_gpu(sc, 3, binning=binning)  # prepare + resume under the same capacity
    for k in ("color", "depth", "alpha", "radii"):
        assert torch.equal(first[k], again[k]), k
    for k in ("keys", "vals", "ranges", "n_contrib", "final_T"):
        assert torch.equal(first["views"][k], again["views"][k]), k
    # the previous capacity is too small for a larger scene of the same size: falls back
    raster._LAST_CAPACITY[key] = 256
    grown = run_gpu(sc, 3, binning=binning)
    for k in ("color", "depth", "radii"):
        assert torch.equal(first[k], grown[k]), k
    assert grown["st"].capacity >= grown["st"].num_rendered > 256


def test_forwards_from_two_threads_do_not_share_the_readback():
    """Each thread reads num_rendered back through a pinned buffer and event
    of its own: two threads rendering different scenes at the same time get
    exactly their single-threaded results."""
    import threading
    from gs_hip import camera as gcam
    from gs_hip import raster
    print("two-thread forwards", flush=True)
    dev = torch.device("cuda:0")
    jobs = []
    for k, (P, W, H) in enumerate([(20_000, 320, 180), (60_000, 240, 160)]):
        sc = scenes.plain_scene(seed=80 + k, P=P, W=W, H=H)
        rs = gcam.raster_settings(sc["cam"].to(dev), sc["bg"].to(dev), 3)
        args = [sc[n].to(dev) for n in ("means", "shs")] + [None, sc["opacities"].to(dev), sc["scales"].to(dev),
                                                            sc["rotations"].to(dev), None]
        jobs.append((rs, args))
    ref = []
    for rs, args in jobs:
        c, r, _, _, st = raster.forward(rs, *args, want_alpha=False)
        ref.append((c.clone(), st.num_rendered))
    torch.cuda.synchronize()
    errors = []

    def work(k):
        try:
            rs, args = jobs[k]
            with torch.cuda.stream(torch.cuda.Stream(device=dev)):
                for _ in range(25):
                    c, _, _, _, st = raster.forward(rs, *args, want_alpha=False)
                    torch.cuda.current_stream(dev).synchronize()
                    if st.num_rendered != ref[k][1] or not torch.equal(c, ref[k][0]):
                        errors.append((k, st.num_rendered))
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append((k, repr(e)))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts)
    assert not errors, errors[:4]
