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
        return self._fused_step(_InBackward, (means, shs))

    def in_backward_activations(self, raw_scale: torch.Tensor, raw_rot: torch.Tensor,
                                raw_opacity: torch.Tensor) -> "_ActInBackward":
        """The same for the raw scaling, rotation and opacity parameters: the
        activations' backward (``gs_hip.model.activate(..., adam=...)``)
        applies their Adam step itself."""
        return self._fused_step(_ActInBackward, (raw_scale, raw_rot, raw_opacity))

    def _fused_step(self, cls, params):
        cache = self.__dict__.setdefault("_fused_cache", {})
        key = (cls, tuple(id(p) for p in params))
        fs = cache.get(key)
        if fs is None or any(a is not b for a, b in zip(fs.params, params)):
            fs = cache[key] = cls(self, params)
        return fs

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


class _FusedStep:
    """Parameters whose Adam step a backward kernel applies.  The checks and
    the argument struct are built once and reused (only the step count and
    the learning rates are refreshed per step) while the parameters' and
    moments' storage and the groups' betas and eps stay the same;
    ``take()`` advances the step counts and returns the struct."""

    def __init__(self, opt: FusedAdam, params):
        self.opt, self.params = opt, tuple(params)
        self._key = None
        self._struct = None

    def _storage_key(self):
        """Storage of the parameters and moments, and the groups' betas and
        eps: a change in any of them rebuilds (and re-checks) the struct."""
        key = []
        for p in self.params:
            st = self.opt.state.get(p)
            g = self.opt._group_of(p)
            key.append((p.data_ptr(), st["exp_avg"].data_ptr() if st else 0, st["exp_avg_sq"].data_ptr() if st else 0,
                        tuple(float(b) for b in g["betas"]), float(g["eps"])))
        return tuple(key)

    def take(self):
        if self._storage_key() != self._key:
            groups, states = _check_fused(self.opt, self.params)
            self._groups, self._states = groups, states
            self._struct = self._build(groups, states)
            self._key = self._storage_key()
        states = self._states
        step = states[0]["step"]
        if any(st["step"] != step for st in states):
            raise RuntimeError("adam in backward: the parameters have taken different numbers of steps")
        for st in states:
            st["step"] = step + 1
        self._refresh(self._struct, self._groups, step + 1)
        return self._struct


class _ActInBackward(_FusedStep):
    """The raw (scaling, rotation, opacity): a gs_adam_act_t."""

    def _build(self, groups, states):
        ss, sr, so = states
        b1, b2 = groups[0]["betas"]
        return _lib.GsAdamAct(scale_m=ss["exp_avg"].data_ptr(), scale_v=ss["exp_avg_sq"].data_ptr(),
                              rot_m=sr["exp_avg"].data_ptr(), rot_v=sr["exp_avg_sq"].data_ptr(),
                              opacity_m=so["exp_avg"].data_ptr(), opacity_v=so["exp_avg_sq"].data_ptr(),
                              beta1=float(b1), beta2=float(b2), eps=float(groups[0]["eps"]))

    @staticmethod
    def _refresh(a, groups, step):
        gs, gr, go = groups
        a.scale_lr, a.rot_lr, a.opacity_lr = float(gs["lr"]), float(gr["lr"]), float(go["lr"])
        a.step = step


class _InBackward(_FusedStep):
    """(means, shs): a gs_adam_fused_t."""

    def _build(self, groups, states):
        gm, gs = groups
        sm, ss = states
        row = int(gs.get("row", 0))
        if row and row != self.params[1][0].numel():
            raise RuntimeError("adam in backward: the SH group's row must be one Gaussian's coefficients")
        b1, b2 = gm["betas"]
        return _lib.GsAdamFused(means_m=sm["exp_avg"].data_ptr(), means_v=sm["exp_avg_sq"].data_ptr(),
                                shs_m=ss["exp_avg"].data_ptr(), shs_v=ss["exp_avg_sq"].data_ptr(),
                                shs_lead=int(gs.get("lead", 0)) if row else 0, beta1=float(b1), beta2=float(b2),
                                eps=float(gm["eps"]))

    @staticmethod
    def _refresh(a, groups, step):
        gm, gs = groups
        a.means_lr, a.shs_lr = float(gm["lr"]), float(gs["lr"])
        a.shs_lead_lr = float(gs.get("lead_lr", gs["lr"])) if gs.get("row", 0) else float(gs["lr"])
        a.step = step
