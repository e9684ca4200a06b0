"""Radix sort and kNN parity on the GPU.

Sort: bit-exact against numpy's stable sort (keys and carried values), so
equal keys keep their input order.  kNN: bit-exact against the oracle's
brute-force O(P^2) search (same squared-distance formula, three best summed
smallest first).
"""
import numpy as np
import pytest
import torch

from gpu_helpers import DEV, np_
from gs_hip import raster
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,end_bit,dup", [(0, 40, 1), (1, 40, 1), (4095, 40, 3), (4097, 43, 7),
                                           (100_000, 64, 1), (300_000, 43, 50), (1_000_003, 45, 1000)])
def test_sort_pairs_stable(n, end_bit, dup):
    rng = np.random.default_rng(n + end_bit)
    tiles = rng.integers(0, max(1, (1 << (end_bit - 32)) if end_bit > 32 else 1), size=n, dtype=np.uint64)
    depth = rng.integers(0, max(1, (1 << 32) // dup), size=n, dtype=np.uint64) * np.uint64(dup)
    keys = (tiles << np.uint64(32)) | depth
    if end_bit < 64:
        keys &= np.uint64((1 << end_bit) - 1)
    vals = np.arange(n, dtype=np.uint32)
    k, v = raster.sort_pairs_u64(torch.from_numpy(keys.view(np.int64)).to(DEV),
                                 torch.from_numpy(vals.view(np.int32)).to(DEV), end_bit)
    torch.cuda.synchronize()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(np_(k).view(np.uint64), keys[order])
    np.testing.assert_array_equal(np_(v).view(np.uint32), vals[order])


def _knn_case(pts):
    pts = np.ascontiguousarray(pts, np.float32)
    got = np_(raster.knn3_meandist(torch.from_numpy(pts).to(DEV)))
    ref = O.knn3_brute(pts)
    np.testing.assert_array_equal(got, ref)


def test_knn_uniform():
    rng = np.random.default_rng(0)
    _knn_case(rng.uniform(-1, 1, (20000, 3)))


def test_knn_clustered_and_duplicates():
    rng = np.random.default_rng(1)
    c = rng.uniform(-3, 3, (16, 3))
    pts = c[rng.integers(0, 16, 15000)] + rng.normal(0, 0.05, (15000, 3))
    pts[100:200] = pts[0]  # exact duplicates -> zero distances
    _knn_case(pts)


def test_knn_surface_points():
    from gs_hip import scenes
    f = scenes.depth_lift_frame(3, W=160, H=90)
    _knn_case(f["xyz"].numpy())


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_knn_tiny(P):
    rng = np.random.default_rng(P)
    _knn_case(rng.normal(0, 1, (P, 3)))


def test_knn_flat_plane():
    rng = np.random.default_rng(7)
    pts = rng.uniform(-1, 1, (8000, 3))
    pts[:, 2] = 0.5
    _knn_case(pts)


def test_distCUDA2_drop_in():
    from simple_knn._C import distCUDA2
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, (5000, 3)).astype(np.float32)
    d = distCUDA2(torch.from_numpy(pts).float().cuda())
    assert d.shape == (5000,) and d.dtype == torch.float32 and d.device.type == "cuda"
    np.testing.assert_array_equal(np_(d), O.knn3_brute(pts))
