"""FusedAdam (one HIP launch) against torch.optim.Adam with the same groups,
learning rates, betas and eps=1e-15, over several steps (rtol 1e-5)."""
import pytest
import torch

from gs_hip.optim import FusedAdam

pytestmark = pytest.mark.gpu


def test_fused_adam_matches_torch():
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(0)
    shapes = [(1001, 3), (1001, 1, 3), (1001, 15, 3), (1001, 1), (1001, 3), (1001, 4), (7,)]
    lrs = [1.6e-4, 2.5e-3, 1.25e-4, 5e-2, 5e-3, 1e-3, 1e-3]
    a = [torch.randn(s, generator=g).to(dev).requires_grad_(True) for s in shapes]
    b = [t.detach().clone().requires_grad_(True) for t in a]
    oa = FusedAdam([{"params": [t], "lr": lr} for t, lr in zip(a, lrs)], lr=0.0, eps=1e-15)
    ob = torch.optim.Adam([{"params": [t], "lr": lr} for t, lr in zip(b, lrs)], lr=0.0, eps=1e-15)
    for it in range(5):
        grads = [torch.randn(s, generator=g).to(dev) for s in shapes]
        for t, gr in zip(a, grads):
            t.grad = gr.clone()
        for t, gr in zip(b, grads):
            t.grad = gr.clone()
        if it == 3:
            a[6].grad = None  # a parameter without gradient is left untouched
            b[6].grad = None
        oa.step()
        ob.step()
    for x, y in zip(a, b):
        torch.testing.assert_close(x, y, rtol=1e-5, atol=1e-7)
