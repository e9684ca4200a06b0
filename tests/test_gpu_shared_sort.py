"""The depth-first binning split around one depth sort shared by many frames
(gs_render_depth_keys, gs_multi_depth_sort, gs_render_after_depth_sort),
and also around one shared list sort (gs_render_binning_duplicate,
gs_multi_list_sort, gs_render_after_list_sort), leaves exactly what
gs_render leaves, frame by frame: the same depth order, the same sorted tile
lists and ranges, the same image and per-pixel state.

Frames of the batched pose workload's kinds: different resolutions and
Gaussian counts (each with its own workspaces), sorted together; their list
sorts take one pass (up to 256 tiles) or two."""
import ctypes

import numpy as np
import pytest
import torch

from gs_hip import _lib
from gs_hip import raster as _raster
from gs_hip.batch import (BinningStepsSet, DepthSortSet, FrameSets, KeysSet, ListSortSet, launch_slot_after_sort,
                          launch_slot_dup, launch_slot_keys, launch_slot_tail, make_slot, start_capacities)
from gs_hip.pose import build_instance

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _u32(t: torch.Tensor, off: int, count: int) -> np.ndarray:
    return t[off:off + 4 * count].cpu().numpy().view(np.uint32)


def _f32(t: torch.Tensor, off: int, count: int) -> np.ndarray:
    return t[off:off + 4 * count].cpu().numpy().view(np.float32)


@pytest.mark.parametrize("keys_set", [False, True])
@pytest.mark.parametrize("shared_list", [False, True])
@pytest.mark.parametrize("sizes", [[(160, 90, 12000), (320, 180, 40000), (128, 72, 5000)],
                                   [(960, 540, 200000), (480, 270, 60000), (320, 180, 20000)]])
def test_shared_sorts_equal_per_frame_binning(sizes, shared_list, keys_set):
    print("shared depth sort, shared list sort", shared_list, "one-set preprocess and keys", keys_set, "frames", sizes,
          flush=True)
    insts = [build_instance(300 + k, W=W, H=H, P=P) for k, (W, H, P) in enumerate(sizes)]
    pose = torch.tensor([0.0, 0, 0, 0, 0, 0, 1], dtype=torch.float32, device=DEV).repeat(len(insts), 1)
    caps = start_capacities(insts, pose, _lib.BINNING_DEPTH_FIRST, 1.3)
    grad = torch.zeros((len(insts), 7), dtype=torch.float32, device=DEV)
    a = [make_slot(i, c, _lib.BINNING_DEPTH_FIRST, pose[k].data_ptr(), grad[k].data_ptr())
         for k, (i, c) in enumerate(zip(insts, caps))]
    b = [make_slot(i, c, _lib.BINNING_DEPTH_FIRST, pose[k].data_ptr(), grad[k].data_ptr())
         for k, (i, c) in enumerate(zip(insts, caps))]
    lib = _lib.load()
    stream = _raster._stream(DEV)
    c = ctypes.c_void_p
    # per frame: preprocess + gs_render (binning with its own depth sort, blend)
    for s in a:
        _lib.check(lib.gs_preprocess(ctypes.byref(s.frame), ctypes.byref(s.gauss), c(s.radii.data_ptr()),
                                     c(s.geom.data_ptr()), s.geom.numel(), c(s.counters.data_ptr()), stream),
                   "gs_preprocess")
        _lib.check(lib.gs_render(ctypes.byref(s.frame), s.inst.P, c(s.geom.data_ptr()), c(s.counters.data_ptr()),
                                 s.capacity, c(s.binning.data_ptr()), s.binning.numel(), c(s.image.data_ptr()),
                                 s.image.numel(), c(s.color.data_ptr()), c(s.depth.data_ptr()), None, stream),
                   "gs_render")
    # the split: preprocess and depth keys per frame (or as one set), one
    # segmented depth sort, the rest per frame
    if keys_set:
        KeysSet(b, DEV).launch(stream)
    else:
        for s in b:
            launch_slot_keys(s, stream)
    DepthSortSet(b, DEV).launch(stream)
    if shared_list and keys_set:
        # with the one-set duplicates and checkpoint scans, and one blend launch
        steps = BinningStepsSet(b, DEV)
        steps.duplicate(stream)
        ListSortSet(b, DEV).launch(stream)
        steps.after_list_sort(stream)
        FrameSets(b, DEV).blend_launch(stream)
    elif shared_list:
        for s in b:
            launch_slot_dup(s, stream)
        ListSortSet(b, DEV).launch(stream)
        for s in b:
            launch_slot_tail(s, stream, blend=True)
    else:
        for s in b:
            launch_slot_after_sort(s, stream, blend=True)
    torch.cuda.synchronize()
    for k, (sa, sb) in enumerate(zip(a, b)):
        W, H, P = sa.inst.W, sa.inst.H, sa.inst.P
        ca, cb = sa.counters.cpu().numpy(), sb.counters.cpu().numpy()
        print(k, "num_rendered", ca[0], "visible", ca[2], "status", ca[3], flush=True)
        assert (ca == cb).all()
        assert ca[3] == 0 and 0 < ca[0] <= sa.capacity
        LG = _raster._geom_layout(P)
        assert torch.equal(sa.radii, sb.radii)
        tiles = _u32(sa.geom, LG.tiles, P)
        assert np.array_equal(tiles, _u32(sb.geom, LG.tiles, P))
        vis = tiles != 0  # (the splat records of culled Gaussians are not written)
        assert np.array_equal(_u32(sa.geom, LG.splat, 12 * P).reshape(P, 12)[vis],
                              _u32(sb.geom, LG.splat, 12 * P).reshape(P, 12)[vis])
        assert np.array_equal(_u32(sa.geom, LG.offsets, P), _u32(sb.geom, LG.offsets, P))
        LB = _raster._binning_layout(sa.capacity, P, W, H, _lib.BINNING_DEPTH_FIRST)
        LI = _raster._image_layout(W, H)
        nt = ((W + 15) // 16) * ((H + 15) // 16)
        # depth order of the visible Gaussians (the result of the depth sort)
        assert np.array_equal(_u32(sa.binning, LB.dvals_a, int(ca[2])), _u32(sb.binning, LB.dvals_a, int(ca[2])))
        assert np.array_equal(_u32(sa.binning, LB.vals_sorted, int(ca[0])),
                              _u32(sb.binning, LB.vals_sorted, int(ca[0])))
        assert np.array_equal(_u32(sa.image, LI.ranges, 2 * nt), _u32(sb.image, LI.ranges, 2 * nt))
        assert np.array_equal(_u32(sa.image, LI.n_contrib, W * H), _u32(sb.image, LI.n_contrib, W * H))
        assert np.array_equal(_f32(sa.image, LI.final_T, W * H), _f32(sb.image, LI.final_T, W * H))
        assert torch.equal(sa.color, sb.color) and torch.equal(sa.depth, sb.depth)
