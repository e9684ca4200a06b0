"""32-bit-key radix sort (used by the depth-first binning) against numpy's
stable sort, keys and carried values bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, np_
from gs_hip import _lib
from gs_hip.raster import _ptr, _stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,end_bit,hi", [(1, 32, 1 << 32), (5000, 32, 1 << 32), (300_001, 11, 2040),
                                          (2_000_000, 13, 8160), (70_000, 1, 1), (3_000_001, 11, 2040),
                                          (700_001, 10, 1024)])
def test_sort_pairs_u32_stable(n, end_bit, hi):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    lib = _lib.load()
    ka = torch.from_numpy(keys.view(np.int32)).to(DEV)
    va = torch.from_numpy(vals.view(np.int32)).to(DEV)
    kb, vb = torch.empty_like(ka), torch.empty_like(va)
    n_dev = torch.tensor([n], dtype=torch.int32, device=DEV)
    tmp = torch.empty(int(lib.gs_sort_tmp_bytes(n)), dtype=torch.uint8, device=DEV)
    in_b = ctypes.c_int(0)
    _lib.check(lib.gs_sort_pairs_u32(_ptr(ka), _ptr(va), _ptr(kb), _ptr(vb), _ptr(n_dev), n, end_bit, _ptr(tmp),
                                     tmp.numel(), ctypes.byref(in_b), _stream(DEV)), "gs_sort_pairs_u32")
    k, v = (kb, vb) if in_b.value else (ka, va)
    torch.cuda.synchronize()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(np_(k).view(np.uint32), keys[order])
    np.testing.assert_array_equal(np_(v).view(np.uint32), vals[order])
