"""gs_knn3_meandist_static: distCUDA2's kernels over the library's static
device scratch instead of a caller workspace (for callers that cannot
allocate).  Called through the C ABI, it must equal gs_knn3_meandist bit for
bit on consecutive calls that reuse the scratch -- a large input, then a
smaller one (stale contents of the first call must not leak into the
second), then the maximum size -- and refuse a larger input before anything
is enqueued."""
import ctypes

import pytest
import torch

from gs_hip import _lib

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
MAX_POINTS = 1 << 20


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)


def _static(points):
    lib = _lib.load()
    out = torch.full((points.shape[0],), float("nan"), device=DEV)
    rc = lib.gs_knn3_meandist_static(ctypes.c_void_p(points.data_ptr()), points.shape[0],
                                     ctypes.c_void_p(out.data_ptr()), _stream())
    assert rc == _lib.GS_OK, lib.gs_error_string(rc)
    return out


def _workspace(points):
    lib = _lib.load()
    P = points.shape[0]
    out = torch.full((P,), float("nan"), device=DEV)
    ws = torch.empty(int(lib.gs_knn_workspace_bytes(P)), dtype=torch.uint8, device=DEV)
    rc = lib.gs_knn3_meandist(ctypes.c_void_p(points.data_ptr()), P, ctypes.c_void_p(out.data_ptr()),
                              ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream())
    assert rc == _lib.GS_OK, lib.gs_error_string(rc)
    return out


def _clustered(n, seed):
    g = torch.Generator().manual_seed(seed)
    centres = torch.rand((16, 3), generator=g) * 6 - 3
    return (centres[torch.randint(0, 16, (n,), generator=g)] + 0.2 * torch.randn((n, 3), generator=g)).to(DEV)


def test_static_scratch_equals_the_workspace_entry_point():
    print("no-alloc kNN: 600k, then 70,001, then the maximum", flush=True)
    g = torch.Generator().manual_seed(0)
    inputs = [(torch.rand((600_000, 3), generator=g) * 4 - 2).to(DEV), _clustered(70_001, 1),
              _clustered(MAX_POINTS, 2)]
    for pts in inputs:
        a = _static(pts)
        b = _workspace(pts)
        torch.cuda.synchronize()
        assert torch.isfinite(a).all()
        assert torch.equal(a, b), pts.shape


def test_static_scratch_refuses_more_than_its_maximum():
    lib = _lib.load()
    pts = torch.zeros((MAX_POINTS + 1, 3), device=DEV)
    out = torch.zeros(MAX_POINTS + 1, device=DEV)
    rc = lib.gs_knn3_meandist_static(ctypes.c_void_p(pts.data_ptr()), MAX_POINTS + 1,
                                     ctypes.c_void_p(out.data_ptr()), _stream())
    assert rc == 4 and b"static scratch" in lib.gs_error_string(rc)
    torch.cuda.synchronize()
    assert not out.any()
