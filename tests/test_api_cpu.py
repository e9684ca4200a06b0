"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/gs_raster.h declares, workspace layouts, the
drop-in packages' argument errors (raised before any device work), the
loud failure on CPU tensors, and the camera conventions."""
import math
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "gs_raster.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gs_hip import _lib
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert lib.gs_abi_version() == 4


def test_layouts():
    from gs_hip import _lib
    L = _lib.binning_layout(1000, 500, 960, 540, _lib.BINNING_SORT64)
    # 8-bit digits: 2040 tiles -> 11 bits -> 43-bit keys -> 6 passes
    assert L.sort_passes == 6 and L.key_bytes == 8
    assert _lib.binning_layout(10, 5, 256, 256, _lib.BINNING_SORT64).sort_passes == 5  # 8 + 32 bits
    assert _lib.binning_layout(10, 5, 16, 16, _lib.BINNING_SORT64).sort_passes == 4  # 1 tile: 32 bits
    D = _lib.binning_layout(1000, 500, 960, 540, _lib.BINNING_DEPTH_FIRST)
    assert D.sort_passes == 2 and D.key_bytes == 4  # tile ids only: 11 bits
    assert _lib.binning_layout(1000, 500, 1920, 1080, _lib.BINNING_DEPTH_FIRST).sort_passes == 2  # 13 bits
    # one word per list entry while bits(P) + bits(tiles) <= 32, ids in their own array otherwise
    Dp = _lib.binning_layout(1000, 500_000, 960, 540, _lib.BINNING_DEPTH_FIRST)
    assert Dp.id_bits == 19 and Dp.vals_sorted == Dp.keys_sorted and Dp.sort_passes == 2  # 19 + 11 bits
    assert _lib.binning_layout(1000, 524_288, 1920, 1080, _lib.BINNING_DEPTH_FIRST).id_bits == 19  # 19 + 13
    assert _lib.binning_layout(1000, 524_289, 1920, 1080, _lib.BINNING_DEPTH_FIRST).id_bits == 0   # 20 + 13
    assert _lib.binning_layout(1000, 500_000, 960, 540, _lib.BINNING_SORT64).id_bits == 0
    assert D.dscan > D.dvals_b > D.dkeys_b > D.dvals_a > D.dkeys_a >= D.vals_alt
    G = _lib.geom_layout(1000)
    assert G.total >= 1000 * (48 + 24 + 8) and G.splat % 256 == 0 and G.cov3D % 256 == 0
    I = _lib.image_layout(250, 131)
    assert I.total >= 8 * 16 * 9 + 8 * 250 * 131
    assert I.tile_order >= I.tile_maxc + 4 * 16 * 9 and I.total >= I.tile_order + 4 * 16 * 9
    with pytest.raises(_lib.GsError):
        _lib.binning_layout(-1, 10, 10, 10)
    with pytest.raises(_lib.GsError):
        _lib.binning_layout(10, 10, 10, 10, 7)


def test_c_abi_rejects_bad_arguments_without_device():
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    f = _lib.GsFrame()
    f.width, f.height = 0, 10  # invalid
    g = _lib.GsGaussians()
    assert lib.gs_preprocess(ctypes.byref(f), ctypes.byref(g), None, None, 0, None, None) == 1
    assert lib.gs_knn3_meandist(None, -1, None, None, 0, None) == 1
    # the no-workspace kNN refuses more points than its static scratch holds
    # before it touches a device (the pointers are never read)
    one = ctypes.c_void_p(16)
    assert lib.gs_knn3_meandist_static(one, (1 << 20) + 1, one, None) == 4
    assert b"static scratch" in lib.gs_error_string(4)
    assert lib.gs_knn3_meandist_static(None, 5, None, None) == 1
    assert lib.gs_error_string(1) == b"invalid argument"
    # the optimiser-in-backward entry points refuse missing arguments before any launch
    ad = _lib.GsAdamAct(step=1)
    assert lib.gs_activate_backward_adam(-1, None, None, None, None, None, None, None, None, ctypes.byref(ad),
                                         None) == 1
    assert lib.gs_activate_backward_adam(10, None, None, None, None, None, None, None, None, None, None) == 1
    assert lib.gs_activate_backward_adam(10, None, None, None, None, None, None, None, None, ctypes.byref(ad),
                                         None) == 1
    assert lib.gs_activate_backward_adam(0, None, None, None, None, None, None, None, None, ctypes.byref(ad),
                                         None) == 0
    gr = _lib.GsGrads()
    fa = _lib.GsAdamFused(step=1)
    assert lib.gs_backward_adam(ctypes.byref(f), ctypes.byref(g), None, None, None, 0, None, None, None, None, None,
                                ctypes.byref(gr), ctypes.byref(fa), None, 0, None) == 1


def _settings():
    from diff_gaussian_rasterization import GaussianRasterizationSettings
    return GaussianRasterizationSettings(
        image_height=32, image_width=32, tanfovx=0.5, tanfovy=0.5, bg=torch.zeros(3), scale_modifier=1.0,
        viewmatrix=torch.eye(4), projmatrix=torch.eye(4), sh_degree=0, campos=torch.zeros(3),
        prefiltered=False, debug=False)


def test_rasterizer_argument_errors():
    from diff_gaussian_rasterization import GaussianRasterizer
    r = GaussianRasterizer(_settings())
    m = torch.zeros(4, 3)
    with pytest.raises(Exception, match="exactly one of either SHs|excatly one of either SHs"):
        r(means3D=m, means2D=m, opacities=torch.ones(4, 1), shs=None, colors_precomp=None,
          scales=torch.ones(4, 3), rotations=torch.ones(4, 4))
    with pytest.raises(Exception, match="SHs or precomputed colors"):
        r(means3D=m, means2D=m, opacities=torch.ones(4, 1), shs=torch.zeros(4, 1, 3), colors_precomp=torch.zeros(4, 3),
          scales=torch.ones(4, 3), rotations=torch.ones(4, 4))
    with pytest.raises(Exception, match="scale/rotation pair or precomputed 3D covariance"):
        r(means3D=m, means2D=m, opacities=torch.ones(4, 1), colors_precomp=torch.zeros(4, 3),
          scales=torch.ones(4, 3), rotations=torch.ones(4, 4), cov3D_precomp=torch.zeros(4, 6))
    with pytest.raises(Exception, match="scale/rotation pair or precomputed 3D covariance"):
        r(means3D=m, means2D=m, opacities=torch.ones(4, 1), colors_precomp=torch.zeros(4, 3))


def test_cpu_tensors_fail_loudly():
    from diff_gaussian_rasterization import GaussianRasterizer
    r = GaussianRasterizer(_settings())
    m = torch.zeros(4, 3)
    with pytest.raises(RuntimeError, match="HIP device"):
        r(means3D=m, means2D=m, opacities=torch.ones(4, 1), colors_precomp=torch.zeros(4, 3),
          scales=torch.ones(4, 3), rotations=torch.ones(4, 4))
    from simple_knn._C import distCUDA2
    with pytest.raises(RuntimeError, match="HIP device"):
        distCUDA2(torch.zeros(10, 3))


def test_camera_matrices_follow_reference_conventions():
    from gs_hip.camera import fov2focal, look_at, make_camera
    W, H = 64, 48
    fovx = math.radians(60)
    fovy = 2 * math.atan(math.tan(fovx / 2) * H / W)
    R, t = look_at([1.0, 2.0, -5.0], [0.0, 0.0, 0.0])
    cam = make_camera(R, t, fovx, fovy, W, H)
    # camera centre recovered from the inverse view matrix
    np.testing.assert_allclose(cam.camera_center.numpy(), [1.0, 2.0, -5.0], atol=1e-5)
    # a point on the optical axis projects to NDC (0, 0) with w = its depth
    p = np.array([0.0, 0.0, 0.0, 1.0], np.float32)
    h = p @ cam.full_proj_transform.numpy()
    np.testing.assert_allclose(h[:2] / h[3], [0.0, 0.0], atol=1e-6)
    assert h[3] == pytest.approx(math.sqrt(1 + 4 + 25), rel=1e-5)
    # x right, y down, z forward: a point to the camera's right has positive NDC x
    right = R[:, 0]
    h2 = np.concatenate([right * 0.5, [1.0]]).astype(np.float32) @ cam.full_proj_transform.numpy()
    assert h2[0] / h2[3] > 0
    assert fov2focal(fovx, W) == pytest.approx(W / (2 * math.tan(fovx / 2)))


def test_loss_matches_reference_formula():
    """The torch SSIM equals a direct per-pixel evaluation of the windowed
    statistics (zero padding) at a few pixels."""
    from gs_hip.loss import gaussian_window, ssim
    g = torch.Generator().manual_seed(0)
    a = torch.rand((3, 20, 24), generator=g)
    b = torch.rand((3, 20, 24), generator=g)
    s = ssim(a, b)
    w = gaussian_window(11, 1.5, 1)[0, 0].double()
    ap = torch.nn.functional.pad(a.double(), (5, 5, 5, 5))
    bp = torch.nn.functional.pad(b.double(), (5, 5, 5, 5))
    tot = 0.0
    for c in range(3):
        for y in range(20):
            for x in range(24):
                pa, pb = ap[c, y:y + 11, x:x + 11], bp[c, y:y + 11, x:x + 11]
                m1, m2 = (w * pa).sum(), (w * pb).sum()
                v1 = (w * pa * pa).sum() - m1 * m1
                v2 = (w * pb * pb).sum() - m2 * m2
                c12 = (w * pa * pb).sum() - m1 * m2
                C1, C2 = 1e-4, 9e-4
                tot += ((2 * m1 * m2 + C1) * (2 * c12 + C2) / ((m1 * m1 + m2 * m2 + C1) * (v1 + v2 + C2))).item()
    assert float(s) == pytest.approx(tot / (3 * 20 * 24), rel=1e-5)


def test_multi_frame_set_builders_on_the_host():
    """The set builders only fill a host buffer: tile / work-item prefixes per
    frame, and argument errors, checked without a device (the pointers are
    never dereferenced on the host)."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    n = 3
    frames = (_lib.GsFrame * n)()
    sizes = [(160, 90), (33, 17), (256, 256)]
    dummy = ctypes.c_float(0.0)
    for f, (W, H) in zip(frames, sizes):
        f.width, f.height = W, H
        f.tanfovx = f.tanfovy = 0.5
        f.scale_modifier = 1.0
        f.viewmatrix = f.projmatrix = f.bg = ctypes.addressof(dummy)
        f.deterministic = 1
    P = (ctypes.c_int * n)(100, 0, 50)
    cap = (ctypes.c_longlong * n)(1000, 0, 70000)
    ptr = (ctypes.c_void_p * n)(*[0x1000] * n)
    tiles = [((W + 15) // 16) * ((H + 15) // 16) for W, H in sizes]
    nb = lib.gs_multi_blend_bytes(n)
    host = (ctypes.c_uint8 * nb)()
    total = ctypes.c_uint(0)
    assert lib.gs_multi_blend_build(n, frames, P, ptr, cap, ptr, ptr, ptr, ptr, None, host, nb,
                                    ctypes.byref(total)) == 0
    prefix = np.frombuffer(bytes(host)[:4 * (n + 1)], dtype=np.uint32)
    assert list(prefix) == [0, tiles[0], tiles[0] + tiles[1], sum(tiles)] and total.value == sum(tiles)
    assert lib.gs_multi_blend_build(n, frames, P, ptr, cap, ptr, ptr, ptr, ptr, None, host, nb - 1,
                                    ctypes.byref(total)) == 2  # workspace too small
    assert lib.gs_multi_blend_build(0, frames, P, ptr, cap, ptr, ptr, ptr, ptr, None, host, nb,
                                    ctypes.byref(total)) == 1
    nr = lib.gs_multi_render_bwd_bytes(n)
    host2 = (ctypes.c_uint8 * nr)()
    blocks = ctypes.c_uint(0)
    assert lib.gs_multi_render_bwd_build(n, frames, P, ptr, cap, ptr, ptr, ptr, None, None, ptr, host2, nr,
                                         ctypes.byref(blocks)) == 0
    # work-item bound per frame: capacity / 256 + tiles + 1 (frames without Gaussians or capacity get none)
    bounds = [cap[0] // 256 + tiles[0] + 1, 0, cap[2] // 256 + tiles[2] + 1]
    prefix2 = np.frombuffer(bytes(host2)[:4 * (n + 1)], dtype=np.uint32)
    assert list(prefix2) == [0, bounds[0], bounds[0], sum(bounds)] and blocks.value == sum(bounds)
    frames[1].deterministic = 0  # one kernel variant per set
    assert lib.gs_multi_render_bwd_build(n, frames, P, ptr, cap, ptr, ptr, ptr, None, None, ptr, host2, nr,
                                         ctypes.byref(blocks)) == 1


def test_batched_solver_group_layout():
    """Instance groups: every instance in exactly one group, costs dealt
    longest first (no group's load above the lightest one's plus its largest
    item), one group over all streams, several groups one stream each, and
    any other stream count refused."""
    from gs_hip import batch
    cost = [9, 1, 4, 3, 5, 2, 2, 7, 6, 8]
    G, members, gstreams, stream_of = batch.group_layout(cost, 3, 1)
    assert G == 1 and sorted(members[0]) == list(range(10)) and gstreams == [[0, 1, 2]]
    loads = [sum(cost[i] for i in range(10) if stream_of[i] == q) for q in range(3)]
    assert sum(loads) == sum(cost) and max(loads) - min(loads) <= max(cost)
    G, members, gstreams, stream_of = batch.group_layout(cost, 4, 4)
    assert G == 4 and gstreams == [[0], [1], [2], [3]]
    assert sorted(i for m in members for i in m) == list(range(10))
    for j, m in enumerate(members):
        assert all(stream_of[i] == j for i in m)
    loads = [sum(cost[i] for i in m) for m in members]
    assert max(loads) - min(loads) <= max(cost)
    assert batch.group_layout(cost[:2], 2, 4)[0] == 2  # no more groups than instances
    with pytest.raises(ValueError):
        batch.group_layout(cost, 3, 2)
    cands = batch.schedule_candidates((0,), (2, 4, 8), (0, 1, 2, 4, 8), (False,))
    assert cands == [(0, 2, 0, False), (0, 2, 1, False), (0, 2, 2, False), (0, 4, 0, False), (0, 4, 1, False),
                     (0, 4, 4, False), (0, 8, 0, False), (0, 8, 1, False), (0, 8, 8, False)]
    # the shared depth sort is a candidate with the depth-first binning only
    df, s64 = batch._lib.BINNING_DEPTH_FIRST, batch._lib.BINNING_SORT64
    cands = batch.schedule_candidates((df, s64), (4,), (0, 4))
    assert cands == [(df, 4, 0, False), (df, 4, 0, True), (df, 4, 4, False), (df, 4, 4, True),
                     (s64, 4, 0, False), (s64, 4, 4, False)]


def test_multi_depth_sort_set_on_the_host():
    """gs_multi_depth_sort_build only lays out a set (the device addresses it
    is given are never read): one range of upsweep / scatter workgroups per
    frame, ceil(P / 1024) rounded up to a multiple of 8, and refusals of what
    the segmented sort cannot take."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    sizes = [(320, 180, 20000), (960, 540, 200000), (480, 270, 1000)]
    n = len(sizes)
    frames = (_lib.GsFrame * n)()
    for k, (W, H, P) in enumerate(sizes):
        f = frames[k]
        f.width, f.height, f.sh_degree = W, H, 0
        f.viewmatrix, f.projmatrix, f.bg = 0x1000, 0x2000, 0x3000
        f.binning = _lib.BINNING_DEPTH_FIRST
    P = (ctypes.c_int * n)(*[p for _, _, p in sizes])
    cap = (ctypes.c_longlong * n)(*[3 * p for _, _, p in sizes])
    bytes_ = [_lib.binning_layout(3 * p, p, W, H, _lib.BINNING_DEPTH_FIRST).total for W, H, p in sizes]
    bins = (ctypes.c_void_p * n)(*[0x10000000 * (k + 1) for k in range(n)])
    nbytes = (ctypes.c_size_t * n)(*bytes_)
    counters = (ctypes.c_void_p * n)(*[0x7000000 + 256 * k for k in range(n)])
    nb = int(lib.gs_multi_depth_sort_bytes(n))
    assert nb > 0 and lib.gs_multi_depth_sort_bytes(0) == 0
    host = (ctypes.c_uint8 * nb)()
    blocks = ctypes.c_uint(0)
    assert lib.gs_multi_depth_sort_build(n, frames, P, cap, bins, nbytes, counters, host, nb,
                                         ctypes.byref(blocks)) == 0
    prefix = np.frombuffer(bytes(host), dtype=np.uint32, count=n + 1)
    per = [-(-(-(-p // 1024)) // 8) * 8 for _, _, p in sizes]
    assert prefix.tolist() == [0] + np.cumsum(per).tolist() and blocks.value == sum(per)
    # refusals: a set buffer too small, a binning workspace too small, the
    # 64-bit key binning (no depth sort), more keys than a segment takes
    assert lib.gs_multi_depth_sort_build(n, frames, P, cap, bins, nbytes, counters, host, nb - 1,
                                         ctypes.byref(blocks)) == 2
    small = (ctypes.c_size_t * n)(*[b - 1 for b in bytes_])
    assert lib.gs_multi_depth_sort_build(n, frames, P, cap, bins, small, counters, host, nb,
                                         ctypes.byref(blocks)) == 2
    frames[1].binning = _lib.BINNING_SORT64
    assert lib.gs_multi_depth_sort_build(n, frames, P, cap, bins, nbytes, counters, host, nb,
                                         ctypes.byref(blocks)) == 1
    frames[1].binning = _lib.BINNING_DEPTH_FIRST
    big = (ctypes.c_int * n)(20000, (1 << 21) + 1, 1000)
    bigcap = (ctypes.c_longlong * n)(60000, 3 << 21, 3000)
    bigbytes = (ctypes.c_size_t * n)(*[_lib.binning_layout(int(c), int(p), W, H, 0).total
                                       for (W, H, _), p, c in zip(sizes, big, bigcap)])
    assert lib.gs_multi_depth_sort_build(n, frames, big, bigcap, bins, bigbytes, counters, host, nb,
                                         ctypes.byref(blocks)) == 1


def test_multi_list_sort_set_on_the_host():
    """gs_multi_list_sort_build: a frame whose tile ids take 8 bits (240
    tiles) sorts its list in one pass, one with 2040 tiles in two; the second
    pass has workgroups only for the latter.  Two-word list entries are
    refused."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    sizes = [(320, 180, 20000, 60000), (960, 540, 200000, 1500000)]
    n = len(sizes)
    frames = (_lib.GsFrame * n)()
    for k, (W, H, _, _) in enumerate(sizes):
        f = frames[k]
        f.width, f.height, f.sh_degree = W, H, 0
        f.viewmatrix, f.projmatrix, f.bg = 0x1000, 0x2000, 0x3000
        f.binning = _lib.BINNING_DEPTH_FIRST
    P = (ctypes.c_int * n)(*[p for _, _, p, _ in sizes])
    cap = (ctypes.c_longlong * n)(*[c for _, _, _, c in sizes])
    layouts = [_lib.binning_layout(c, p, W, H, _lib.BINNING_DEPTH_FIRST) for W, H, p, c in sizes]
    assert all(L.id_bits > 0 for L in layouts)
    bins = (ctypes.c_void_p * n)(*[0x10000000 * (k + 1) for k in range(n)])
    nbytes = (ctypes.c_size_t * n)(*[L.total for L in layouts])
    counters = (ctypes.c_void_p * n)(*[0x7000000 + 256 * k for k in range(n)])
    images = (ctypes.c_void_p * n)(*[0x9000000 + 0x100000 * k for k in range(n)])
    nb = int(lib.gs_multi_list_sort_bytes(n))
    host = (ctypes.c_uint8 * nb)()
    blocks = (ctypes.c_uint * 4)()
    assert lib.gs_multi_list_sort_build(n, frames, P, cap, bins, nbytes, counters, images, host, nb, blocks) == 0
    per = [-(-(-(-c // 1024)) // 8) * 8 for _, _, _, c in sizes]
    assert list(blocks) == [per[0] + per[1], per[1], 0, 0]
    # two-word entries (ids and tile ids do not fit one word) are refused
    big = (ctypes.c_int * n)(20000, (1 << 21) + 1)
    bigbytes = (ctypes.c_size_t * n)(*[_lib.binning_layout(int(c), int(p), W, H, 0).total
                                       for (W, H, _, c), p in zip(sizes, big)])
    assert _lib.binning_layout(1500000, (1 << 21) + 1, 960, 540, 0).id_bits == 0
    assert lib.gs_multi_list_sort_build(n, frames, big, cap, bins, bigbytes, counters, images, host, nb,
                                        blocks) == 1


def test_shared_stage_entry_points_refuse_bad_arguments_without_device():
    """The split binning, the shared sorts and the multi-image loss check
    their arguments on the host before anything is enqueued (the device
    addresses below are never read)."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    f = _lib.GsFrame()
    f.width, f.height, f.sh_degree = 320, 180, 0
    f.viewmatrix, f.projmatrix, f.bg = 0x1000, 0x2000, 0x3000
    P, cap = 20000, 60000
    f.binning = _lib.BINNING_SORT64
    LB = _lib.binning_layout(cap, P, 320, 180, _lib.BINNING_SORT64)
    LI = _lib.image_layout(320, 180)
    dev = ctypes.c_void_p(0x10000000)
    args = (ctypes.byref(f), P, dev, dev, cap, dev, LB.total, dev, LI.total, None)
    # the phases split around the sorts exist only in the depth-first schedule
    for fn in (lib.gs_render_depth_keys, lib.gs_render_binning_duplicate, lib.gs_render_binning_after_depth_sort,
               lib.gs_render_binning_after_list_sort):
        assert fn(*args) == 1
    f.binning = _lib.BINNING_DEPTH_FIRST
    LB = _lib.binning_layout(cap, P, 320, 180, _lib.BINNING_DEPTH_FIRST)
    short = (ctypes.byref(f), P, dev, dev, cap, dev, LB.total - 1, dev, LI.total, None)
    assert lib.gs_render_depth_keys(*short) == 2
    # launches of empty or missing sets
    assert lib.gs_multi_depth_sort(0, dev, 8, None) == 1
    assert lib.gs_multi_depth_sort(1, None, 8, None) == 1
    assert lib.gs_multi_list_sort(1, dev, None, None) == 1
    zero = (ctypes.c_uint * 4)(0, 0, 0, 0)
    assert lib.gs_multi_list_sort(1, dev, zero, None) == 1
    assert lib.gs_multi_loss(0, dev, 8, None) == 1
    assert lib.gs_multi_loss_bytes(0) == 0 and lib.gs_multi_loss_bytes(2) > 0
    # a loss workspace smaller than gs_loss_workspace_bytes
    n = 1
    vp = (ctypes.c_void_p * n)(0x20000000)
    ints = lambda v: (ctypes.c_int * n)(v)  # noqa: E731
    taps = (ctypes.c_float * 11)(*([1.0 / 11] * 11))
    nb = lib.gs_loss_workspace_bytes(3, 90, 160)
    host = (ctypes.c_uint8 * int(lib.gs_multi_loss_bytes(n)))()
    blocks = ctypes.c_uint(0)
    assert lib.gs_multi_loss_build(n, vp, vp, ints(3), ints(90), ints(160), taps, 0.2, 1, vp, dev, vp,
                                   (ctypes.c_size_t * n)(nb - 1), vp, host, len(host), ctypes.byref(blocks)) == 2
    assert lib.gs_multi_loss_build(n, vp, vp, ints(3), ints(90), ints(160), taps, 0.2, 1, vp, dev, vp,
                                   (ctypes.c_size_t * n)(nb), vp, host, len(host), ctypes.byref(blocks)) == 0
    assert blocks.value == 3 * ((160 + 31) // 32) * ((90 + 15) // 16)  # 32 x 16-pixel tiles per channel


def test_multi_pose_bwd_set_on_the_host():
    """gs_multi_pose_bwd_build lays out, per frame, the workgroups of its own
    backward set-up and one per 256 Gaussians for the finish, and refuses what
    gs_backward_prepare / _finish refuse, plus any frame that does not take
    the pose-only backward.  Built on the host only: the device addresses
    below are never read."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    n = 3
    Ps = [1000, 256, 0]
    caps_l = [4000, 3_000_000, 0]
    frames = (_lib.GsFrame * n)()
    gauss = (_lib.GsGaussians * n)()
    grads = (_lib.GsGrads * n)()
    for i in range(n):
        f = frames[i]
        f.width, f.height, f.sh_degree, f.deterministic = 160, 90, 0, 1
        f.viewmatrix, f.projmatrix, f.bg, f.pose = 0x1000, 0x2000, 0x3000, 0x4000
        f.binning = _lib.BINNING_DEPTH_FIRST
        g = gauss[i]
        g.P, g.M = Ps[i], 1
        g.means3D, g.colors_precomp, g.opacities = 0x5000, 0x6000, 0x7000
        g.scales, g.rotations = 0x8000, 0x9000
        grads[i].dL_dpose = 0xA000
    vp = lambda: (ctypes.c_void_p * n)(*([0x10000000] * n))  # noqa: E731
    caps = (ctypes.c_longlong * n)(*caps_l)
    sb = [int(lib.gs_backward_scratch_bytes(P, c, 1)) for P, c in zip(Ps, caps_l)]
    sbytes = (ctypes.c_size_t * n)(*sb)
    nb = int(lib.gs_multi_pose_bwd_bytes(n))
    assert lib.gs_multi_pose_bwd_bytes(0) == 0 and nb > 0
    host = (ctypes.c_uint8 * nb)()
    blocks = (ctypes.c_uint * 2)()

    def build(nbytes=nb):
        return lib.gs_multi_pose_bwd_build(n, frames, gauss, vp(), vp(), vp(), caps, vp(), vp(), vp(), grads, vp(),
                                           sbytes, host, nbytes, blocks)

    assert build() == 0
    # set-up: the item-table workgroup plus ceil(5 capacity / 1024) clearing
    # ones, at most 2048; none for a frame with no pairs
    prep = [1 + 20, 1 + 2048, 0]
    assert list(blocks) == [sum(prep), 4 + 1 + 0]
    prefix = list((ctypes.c_uint32 * (n + 1)).from_buffer(host))
    assert prefix == [0, 4, 5, 5]
    off = -(-4 * (n + 1) // 256) * 256
    pprefix = list((ctypes.c_uint32 * (n + 1)).from_buffer(host, off))
    assert pprefix == [0, prep[0], prep[0] + prep[1], sum(prep)]
    assert build(nb - 1) == 2
    # a per-Gaussian output: the full backward, not this set's
    grads[1].dL_dmeans3D = 0xB000
    assert build() == 1
    grads[1].dL_dmeans3D = None
    # atomics-order reductions are never the pose-only backward
    frames[0].deterministic = 0
    assert build() == 1
    frames[0].deterministic = 1
    # scratch smaller than gs_backward_scratch_bytes
    sbytes[1] = sb[1] - 1
    assert build() == 2
    sbytes[1] = sb[1]
    assert build() == 0
    dev = ctypes.c_void_p(0x10000000)
    for fn in (lib.gs_multi_pose_bwd_prepare, lib.gs_multi_pose_bwd_finish):
        assert fn(0, dev, blocks, None) == 1
        assert fn(1, None, blocks, None) == 1
        assert fn(1, dev, None, None) == 1


def test_multi_preprocess_keys_set_on_the_host():
    """gs_multi_preprocess_keys_build lays out each frame's preprocess
    workgroups (one per 256 Gaussians) and depth-key workgroups (enough for
    the Gaussians and for the tile ranges it clears), and refuses what
    gs_preprocess / gs_render_depth_keys refuse.  Built on the host only: the
    device addresses below are never read."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    n = 2
    dims = [(160, 90, 1000), (960, 540, 300)]  # 1000 Gaussians; 300 Gaussians over 60 x 34 tiles
    frames = (_lib.GsFrame * n)()
    gauss = (_lib.GsGaussians * n)()
    for i, (W, H, P) in enumerate(dims):
        f = frames[i]
        f.width, f.height, f.sh_degree = W, H, 0
        f.viewmatrix, f.projmatrix, f.bg = 0x1000, 0x2000, 0x3000
        f.binning = _lib.BINNING_DEPTH_FIRST
        g = gauss[i]
        g.P, g.M = P, 1
        g.means3D, g.colors_precomp, g.opacities = 0x5000, 0x6000, 0x7000
        g.scales, g.rotations = 0x8000, 0x9000
    cap = 50000
    vp = lambda: (ctypes.c_void_p * n)(*([0x10000000] * n))  # noqa: E731
    caps = (ctypes.c_longlong * n)(*([cap] * n))
    sz = ctypes.c_size_t
    gb = (sz * n)(*[_lib.geom_layout(P).total for _, _, P in dims])
    bb = (sz * n)(*[_lib.binning_layout(cap, P, W, H, _lib.BINNING_DEPTH_FIRST).total for W, H, P in dims])
    ib = (sz * n)(*[_lib.image_layout(W, H).total for W, H, _ in dims])
    nb = int(lib.gs_multi_preprocess_keys_bytes(n))
    assert lib.gs_multi_preprocess_keys_bytes(0) == 0 and nb > 0
    host = (ctypes.c_uint8 * nb)()
    launch = (ctypes.c_uint * 3)()

    def build(nbytes=nb):
        return lib.gs_multi_preprocess_keys_build(n, frames, gauss, vp(), vp(), gb, vp(), caps, vp(), bb, vp(), ib,
                                                  host, nbytes, launch)

    assert build() == 0
    # preprocess: 4 + 2 workgroups; keys: 4, and 16 for the 2 x 2040 range words
    # of the larger frame; colours precomputed: no SH rows in LDS
    assert list(launch) == [6, 4 + 16, 0]
    pprefix = list((ctypes.c_uint32 * (n + 1)).from_buffer(host))
    kprefix = list((ctypes.c_uint32 * (n + 1)).from_buffer(host, 256))
    assert pprefix == [0, 4, 6] and kprefix == [0, 4, 20]
    assert build(nb - 1) == 2
    # SH coefficients: LDS rows of (3 * 16) | 1 floats for degree 3
    gauss[0].colors_precomp, gauss[0].shs, gauss[0].M = None, 0xA000, 16
    frames[0].sh_degree, frames[0].campos = 3, 0xB000
    assert build() == 0 and launch[2] == 256 * 49 * 4
    frames[0].sh_degree = 4  # beyond degree 3
    assert build() == 1
    frames[0].sh_degree = 3
    # the 64-bit-key schedule has no depth keys
    frames[1].binning = _lib.BINNING_SORT64
    assert build() == 1
    frames[1].binning = _lib.BINNING_DEPTH_FIRST
    # a binning workspace smaller than its layout
    bb[1] -= 1
    assert build() == 2
    bb[1] += 1
    assert build() == 0
    dev = ctypes.c_void_p(0x10000000)
    assert lib.gs_multi_preprocess_keys(0, dev, launch, None) == 1
    assert lib.gs_multi_preprocess_keys(1, None, launch, None) == 1
    assert lib.gs_multi_preprocess_keys(1, dev, None, None) == 1


def test_multi_binning_steps_set_on_the_host():
    """gs_multi_binning_steps_build lays out one workgroup per 256 Gaussians
    of each frame and refuses what gs_render_binning_duplicate refuses.  Built
    on the host only: the device addresses below are never read."""
    import ctypes
    from gs_hip import _lib
    lib = _lib.load()
    n = 3
    dims = [(160, 90, 1000), (960, 540, 300), (320, 180, 513)]
    frames = (_lib.GsFrame * n)()
    for i, (W, H, _) in enumerate(dims):
        f = frames[i]
        f.width, f.height, f.sh_degree = W, H, 0
        f.viewmatrix, f.projmatrix, f.bg = 0x1000, 0x2000, 0x3000
        f.binning = _lib.BINNING_DEPTH_FIRST
    cap = 50000
    P = (ctypes.c_int * n)(*[p for _, _, p in dims])
    vp = lambda: (ctypes.c_void_p * n)(*([0x10000000] * n))  # noqa: E731
    caps = (ctypes.c_longlong * n)(*([cap] * n))
    sz = ctypes.c_size_t
    bb = (sz * n)(*[_lib.binning_layout(cap, p, W, H, _lib.BINNING_DEPTH_FIRST).total for W, H, p in dims])
    ib = (sz * n)(*[_lib.image_layout(W, H).total for W, H, _ in dims])
    nb = int(lib.gs_multi_binning_steps_bytes(n))
    assert lib.gs_multi_binning_steps_bytes(0) == 0 and nb > 0
    host = (ctypes.c_uint8 * nb)()
    blocks = ctypes.c_uint(0)

    def build(nbytes=nb):
        return lib.gs_multi_binning_steps_build(n, frames, P, vp(), vp(), caps, vp(), bb, vp(), ib, host, nbytes,
                                                ctypes.byref(blocks))

    assert build() == 0
    assert blocks.value == 4 + 2 + 3
    assert list((ctypes.c_uint32 * (n + 1)).from_buffer(host)) == [0, 4, 6, 9]
    assert build(nb - 1) == 2
    frames[2].binning = _lib.BINNING_SORT64  # no depth sort to split around
    assert build() == 1
    frames[2].binning = _lib.BINNING_DEPTH_FIRST
    P[1] = 0  # an empty frame has no depth-first steps
    assert build() == 1
    P[1] = 300
    ib[0] -= 1
    assert build() == 2
    ib[0] += 1
    assert build() == 0
    dev = ctypes.c_void_p(0x10000000)
    assert lib.gs_multi_binning_duplicate(0, dev, 1, None) == 1
    assert lib.gs_multi_binning_duplicate(1, None, 1, None) == 1
    assert lib.gs_multi_binning_after_list_sort(0, dev, None) == 1
    assert lib.gs_multi_binning_after_list_sort(1, None, None) == 1
