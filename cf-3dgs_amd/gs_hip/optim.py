"""Adam for the Gaussian parameters in one HIP launch per step.

Same update and defaults as the torch.optim.Adam that CFGaussianModel builds
(scene/gaussian_model_cf.py:256-268: per-group learning rates, betas
(0.9, 0.999), eps 1e-15), with the same ``param_groups`` / ``zero_grad``
surface so learning-rate schedules that edit ``param_groups`` keep working.
Parameters without a gradient are skipped, as torch does.

A group may also carry ``lead_lr``, ``lead`` and ``row``: element i of its
tensors then steps with ``lead_lr`` when ``i % row < lead`` and with ``lr``
otherwise.  ``GaussianParams`` keeps all SH coefficients in one [P,16,3]
tensor and uses this for the reference's f_dc (the leading 3 floats of every
48, lr feature_lr) and f_rest (feature_lr / 20) groups (:257-260).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class FusedAdam(torch.optim.Optimizer):
    def __init__(self, params, lr=0.0, betas=(0.9, 0.999), eps=1e-15):
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))
        self._steps = 0

    def _group_of(self, p):
        for group in self.param_groups:
            if any(q is p for q in group["params"]):
                return group
        raise ValueError("parameter is not managed by this optimiser")

    def _state_of(self, p):
        st = self.state[p]
        if not st:
            st["step"] = 0
            st["exp_avg"] = torch.zeros_like(p)
            st["exp_avg_sq"] = torch.zeros_like(p)
        return st

    def in_backward(self, means: torch.Tensor, shs: torch.Tensor) -> "_InBackward":
        """The Adam step of ``means`` [P,3] and ``shs`` [P,M,3] for the
        rasteriser's backward to apply itself (``adam_in_backward=``): the
        backward then never writes their gradients.  The update is this
        optimiser's, bit for bit; ``step()`` afterwards skips both, since they
        have no gradient."""
        return _InBackward(self, means, shs)

    def in_backward_activations(self, raw_scale: torch.Tensor, raw_rot: torch.Tensor,
                                raw_opacity: torch.Tensor) -> "_ActInBackward":
        """The same for the raw scaling, rotation and opacity parameters: the
        activations' backward (``gs_hip.model.activate(..., adam=...)``)
        applies their Adam step itself."""
        return _ActInBackward(self, raw_scale, raw_rot, raw_opacity)

    @torch.no_grad()
    def step(self, closure=None):
        loss = closure() if closure is not None else None
        lib = _lib.load()
        by_hyper = {}
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.device.type != "cuda" or p.dtype != torch.float32 or not p.is_contiguous():
                    raise RuntimeError("FusedAdam needs contiguous float32 parameters on the GPU")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                g = p.grad if p.grad.is_contiguous() else p.grad.contiguous()
                if g.data_ptr() % 16:
                    g = g.clone()  # the kernel reads 16-byte vectors
                if p.data_ptr() % 16:
                    raise RuntimeError("FusedAdam needs 16-byte aligned parameters (a fresh tensor, not an offset view)")
                key = (float(b1), float(b2), float(group["eps"]), int(st["step"]), p.device)
                row = int(group.get("row", 0))
                lead = int(group.get("lead", 0)) if row else 0
                lead_lr = float(group.get("lead_lr", group["lr"])) if row else 0.0
                by_hyper.setdefault(key, []).append((p, g, st["exp_avg"], st["exp_avg_sq"], float(group["lr"]),
                                                     lead_lr, lead, row))
        for (b1, b2, eps, step, dev), items in by_hyper.items():
            for i in range(0, len(items), 16):
                chunk = items[i:i + 16]
                n = len(chunk)
                P = (ctypes.c_void_p * n)(*[c[0].data_ptr() for c in chunk])
                G = (ctypes.c_void_p * n)(*[c[1].data_ptr() for c in chunk])
                M = (ctypes.c_void_p * n)(*[c[2].data_ptr() for c in chunk])
                V = (ctypes.c_void_p * n)(*[c[3].data_ptr() for c in chunk])
                N = (ctypes.c_longlong * n)(*[c[0].numel() for c in chunk])
                LR = (ctypes.c_float * n)(*[c[4] for c in chunk])
                LLR = (ctypes.c_float * n)(*[c[5] for c in chunk])
                LEAD = (ctypes.c_int * n)(*[c[6] for c in chunk])
                ROW = (ctypes.c_int * n)(*[c[7] for c in chunk])
                stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
                _lib.check(lib.gs_adam_step_rows(n, P, G, M, V, N, LR, LLR, LEAD, ROW, b1, b2, eps, step, stream),
                           "gs_adam_step_rows")
        return loss


def _check_fused(opt: FusedAdam, params) -> tuple:
    """(groups, states) of parameters stepped together inside a backward:
    contiguous aligned float32 GPU tensors, one betas / eps, one step count."""
    groups = [opt._group_of(p) for p in params]
    for p in params:
        if p.device.type != "cuda" or p.dtype != torch.float32 or not p.is_contiguous() or p.data_ptr() % 16:
            raise RuntimeError("adam in backward needs contiguous, 16-byte aligned float32 GPU parameters")
    if len({(tuple(g["betas"]), g["eps"]) for g in groups}) != 1:
        raise RuntimeError("adam in backward: the parameters must share betas and eps")
    states = [opt._state_of(p) for p in params]
    if len({st["step"] for st in states}) != 1:
        raise RuntimeError("adam in backward: the parameters have taken different numbers of steps")
    return groups, states


class _ActInBackward:
    """One pending fused step of the raw (scaling, rotation, opacity);
    ``take()`` advances their step counts and returns the gs_adam_act_t
    fields."""

    def __init__(self, opt: FusedAdam, raw_scale, raw_rot, raw_opacity):
        self.opt, self.params = opt, (raw_scale, raw_rot, raw_opacity)

    def take(self) -> dict:
        (gs, gr, go), states = _check_fused(self.opt, self.params)
        for st in states:
            st["step"] += 1
        ss, sr, so = states
        b1, b2 = gs["betas"]
        return dict(scale_m=ss["exp_avg"].data_ptr(), scale_v=ss["exp_avg_sq"].data_ptr(), scale_lr=float(gs["lr"]),
                    rot_m=sr["exp_avg"].data_ptr(), rot_v=sr["exp_avg_sq"].data_ptr(), rot_lr=float(gr["lr"]),
                    opacity_m=so["exp_avg"].data_ptr(), opacity_v=so["exp_avg_sq"].data_ptr(),
                    opacity_lr=float(go["lr"]), beta1=float(b1), beta2=float(b2), eps=float(gs["eps"]),
                    step=int(ss["step"]))


class _InBackward:
    """One pending fused step of (means, shs); ``take()`` advances their step
    counts and returns the gs_adam_fused_t fields."""

    def __init__(self, opt: FusedAdam, means: torch.Tensor, shs: torch.Tensor):
        self.opt, self.means, self.shs = opt, means, shs

    def take(self) -> dict:
        (gm, gs), (sm, ss) = _check_fused(self.opt, (self.means, self.shs))
        sm["step"] += 1
        ss["step"] += 1
        row = int(gs.get("row", 0))
        lead = int(gs.get("lead", 0)) if row else 0
        if row and row != self.shs[0].numel():
            raise RuntimeError("adam in backward: the SH group's row must be one Gaussian's coefficients")
        b1, b2 = gm["betas"]
        return dict(means_m=sm["exp_avg"].data_ptr(), means_v=sm["exp_avg_sq"].data_ptr(), means_lr=float(gm["lr"]),
                    shs_m=ss["exp_avg"].data_ptr(), shs_v=ss["exp_avg_sq"].data_ptr(), shs_lr=float(gs["lr"]),
                    shs_lead_lr=float(gs.get("lead_lr", gs["lr"])) if row else float(gs["lr"]),
                    shs_lead=lead, beta1=float(b1), beta2=float(b2), eps=float(gm["eps"]), step=int(sm["step"]))
