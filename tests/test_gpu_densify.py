"""Adaptive density control on the GPU against the reference's torch
composition.

The reference (scene/gaussian_model_cf.py:523-719; statistics at
trainer/cf3dgs_trainer.py:144-152) is restated below in the same torch ops --
boolean masks, ``torch.cat`` of the new points and zero moments, boolean
indexing of every parameter and moment -- over separate f_dc / f_rest tensors
and a per-parameter (exp_avg, exp_avg_sq) state, and run on the same inputs
with the same random generator state.  The library's stable compaction must
give the same rows in the same order: every parameter and moment bit for bit,
except the split copies' means, whose rotation goes through torch.bmm here
(a library GEMM) and through an explicit 3x3 product in the kernel (1e-6).
"""
import pytest
import torch
import torch.nn.functional as F

from gs_hip import scenes
from gs_hip.model import GaussianParams

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def build_rotation(r):
    """utils/general_utils.py:76-97."""
    norm = torch.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2] + r[:, 3] * r[:, 3])
    q = r / norm[:, None]
    R = torch.zeros((q.size(0), 3, 3), device=r.device)
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - r * z)
    R[:, 0, 2] = 2 * (x * z + r * y)
    R[:, 1, 0] = 2 * (x * y + r * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - r * x)
    R[:, 2, 0] = 2 * (x * z - r * y)
    R[:, 2, 1] = 2 * (y * z + r * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


class RefModel:
    """The reference's tensors and optimiser state, with its densification
    methods in its torch ops."""
    NAMES = ("xyz", "f_dc", "f_rest", "opacity", "scaling", "rotation")

    def __init__(self, g: GaussianParams):
        f = g._features.detach()
        self.t = {"xyz": g._xyz.detach().clone(), "f_dc": f[:, :1].clone(), "f_rest": f[:, 1:].clone(),
                  "opacity": g._opacity.detach().clone(), "scaling": g._scaling.detach().clone(),
                  "rotation": g._rotation.detach().clone()}
        self.state = {}
        st = g.optimizer.state
        for n, p, sl in (("xyz", g._xyz, None), ("f_dc", g._features, slice(0, 1)), ("f_rest", g._features, slice(1, None)),
                         ("opacity", g._opacity, None), ("scaling", g._scaling, None), ("rotation", g._rotation, None)):
            s = st.get(p)
            if s:
                m, v = s["exp_avg"], s["exp_avg_sq"]
                if sl is not None:
                    m, v = m[:, sl], v[:, sl]
                self.state[n] = [m.clone(), v.clone()]
        self.accum = g.xyz_gradient_accum.clone()
        self.denom = g.denom.clone()
        self.max_radii2D = g.max_radii2D.clone()
        self.percent_dense = g.percent_dense

    def get_scaling(self):
        return torch.exp(self.t["scaling"])

    def get_opacity(self):
        return torch.sigmoid(self.t["opacity"])

    def _prune(self, valid):
        for n in self.NAMES:
            if n in self.state:
                self.state[n] = [self.state[n][0][valid], self.state[n][1][valid]]
            self.t[n] = self.t[n][valid]
        self.accum, self.denom, self.max_radii2D = self.accum[valid], self.denom[valid], self.max_radii2D[valid]

    def _postfix(self, d):
        for n in self.NAMES:
            if n in self.state:
                self.state[n] = [torch.cat((self.state[n][0], torch.zeros_like(d[n])), 0),
                                 torch.cat((self.state[n][1], torch.zeros_like(d[n])), 0)]
            self.t[n] = torch.cat((self.t[n], d[n]), 0)
        P = self.t["xyz"].shape[0]
        self.accum = torch.zeros((P, 1), device=DEV)
        self.denom = torch.zeros((P, 1), device=DEV)
        self.max_radii2D = torch.zeros(P, device=DEV)

    def split(self, grads, thr, extent, N=2):
        n0 = self.t["xyz"].shape[0]
        padded = torch.zeros((n0,), device=DEV)
        padded[:grads.shape[0]] = grads.squeeze()
        sel = torch.where(padded >= thr, True, False)
        sel = torch.logical_and(sel, torch.max(self.get_scaling(), dim=1).values > self.percent_dense * extent)
        stds = self.get_scaling()[sel].repeat(N, 1)
        means = torch.zeros((stds.size(0), 3), device=DEV)
        samples = torch.normal(mean=means, std=stds)
        rots = build_rotation(self.t["rotation"][sel]).repeat(N, 1, 1)
        new_xyz = torch.bmm(rots, samples.unsqueeze(-1)).squeeze(-1) + self.t["xyz"][sel].repeat(N, 1)
        new_scaling = torch.log(self.get_scaling()[sel].repeat(N, 1) / (0.8 * N))
        d = {"xyz": new_xyz, "f_dc": self.t["f_dc"][sel].repeat(N, 1, 1), "f_rest": self.t["f_rest"][sel].repeat(N, 1, 1),
             "opacity": self.t["opacity"][sel].repeat(N, 1), "scaling": new_scaling,
             "rotation": self.t["rotation"][sel].repeat(N, 1)}
        self._postfix(d)
        self._prune(~torch.cat((sel, torch.zeros(N * sel.sum(), device=DEV, dtype=bool))))
        return int(sel.sum())

    def clone(self, grads, thr, extent):
        sel = torch.where(torch.norm(grads, dim=-1) >= thr, True, False)
        sel = torch.logical_and(sel, torch.max(self.get_scaling(), dim=1).values <= self.percent_dense * extent)
        self._postfix({n: self.t[n][sel] for n in self.NAMES})

    def densify_and_prune(self, thr, min_opacity, extent, max_screen_size):
        grads = self.accum / self.denom
        grads[grads.isnan()] = 0.0
        self.clone(grads, thr, extent)
        self.split(grads, thr, extent)
        prune_mask = (self.get_opacity() < min_opacity).squeeze()
        if max_screen_size:
            big_vs = self.max_radii2D > max_screen_size
            big_ws = self.get_scaling().max(dim=1).values > 0.1 * extent
            prune_mask = torch.logical_or(torch.logical_or(prune_mask, big_vs), big_ws)
        self._prune(~prune_mask)

    def prune(self, min_opacity, extent, max_screen_size):
        prune_mask = (self.get_opacity() < min_opacity).squeeze()
        if max_screen_size:
            big_vs = self.max_radii2D > max_screen_size
            big_ws = self.get_scaling().max(dim=1).values > 0.1 * extent
            prune_mask = torch.logical_or(torch.logical_or(prune_mask, big_vs), big_ws)
        self._prune(~prune_mask)


def _trained(P=40_000, seed=3, steps=3):
    """A model after a few training steps with the densification statistics
    gathered the reference's way and with this library's kernel."""
    from gs_hip.loss import fused_photometric_loss
    from gs_hip.model import render
    sc = scenes.training_scene(seed=seed, P=P, W=320, H=180)
    g = GaussianParams(sc["xyz"], sc["features_dc"], sc["features_rest"], sc["scaling"], sc["rotation"],
                       sc["opacity"], sc["sh_degree"], device=DEV)
    g.training_setup()
    cam, bg, target = sc["cam"].to(DEV), sc["bg"].to(DEV), sc["target"].to(DEV)
    ref_accum = torch.zeros((P, 1), device=DEV)
    ref_denom = torch.zeros((P, 1), device=DEV)
    ref_max = torch.zeros(P, device=DEV)
    for _ in range(steps):
        out = render(g, cam, bg, clamp=False)
        loss, _ = fused_photometric_loss(out["image"], target, clamp01=True)
        loss.backward()
        vis, radii, vp = out["visibility_filter"], out["radii"], out["viewspace_points"]
        ref_max[vis] = torch.max(ref_max[vis], radii[vis])
        ref_accum[vis] += torch.norm(vp.grad[vis, :2], dim=-1, keepdim=True)
        ref_denom[vis] += 1
        g.add_densification_stats(vp, radii)
        g.optimizer.step()
        g.optimizer.zero_grad(set_to_none=True)
    torch.cuda.synchronize()
    return g, (ref_accum, ref_denom, ref_max)


def test_densification_stats_equal_the_reference_ops():
    g, (acc, den, mx) = _trained()
    assert torch.equal(g.denom, den) and torch.equal(g.max_radii2D, mx)
    assert int((den > 0).sum()) > 1000
    assert torch.equal(g.xyz_gradient_accum, acc)


def _compare(g: GaussianParams, ref: RefModel):
    f = g._features.detach()
    got = {"xyz": g._xyz.detach(), "f_dc": f[:, :1], "f_rest": f[:, 1:], "opacity": g._opacity.detach(),
           "scaling": g._scaling.detach(), "rotation": g._rotation.detach()}
    for n in RefModel.NAMES:
        assert got[n].shape == ref.t[n].shape, (n, got[n].shape, ref.t[n].shape)
        if n == "xyz":
            torch.testing.assert_close(got[n], ref.t[n], rtol=1e-6, atol=1e-6)
        else:
            assert torch.equal(got[n], ref.t[n]), n
    st = g.optimizer.state
    for n, p, sl in (("xyz", g._xyz, None), ("f_dc", g._features, slice(0, 1)), ("f_rest", g._features, slice(1, None)),
                     ("opacity", g._opacity, None), ("scaling", g._scaling, None), ("rotation", g._rotation, None)):
        m, v = st[p]["exp_avg"], st[p]["exp_avg_sq"]
        if sl is not None:
            m, v = m[:, sl], v[:, sl]
        assert torch.equal(m, ref.state[n][0]) and torch.equal(v, ref.state[n][1]), n
    assert torch.equal(g.denom, ref.denom) and torch.equal(g.max_radii2D, ref.max_radii2D)
    assert torch.equal(g.xyz_gradient_accum, ref.accum)


@pytest.mark.parametrize("max_screen_size", [None, 20])
def test_densify_and_prune_equals_the_reference(max_screen_size):
    print("densify_and_prune, max_screen_size", max_screen_size, flush=True)
    g, _ = _trained()
    ref = RefModel(g)
    # thresholds that select a few percent of the Gaussians each way
    grads = (g.xyz_gradient_accum / g.denom).nan_to_num(0.0)
    thr = float(torch.quantile(grads[grads > 0], 0.9))
    extent = float(torch.exp(g._scaling.detach()).max(dim=1).values.median()) / 0.01
    min_op = float(torch.quantile(torch.sigmoid(g._opacity.detach()), 0.05))
    torch.cuda.manual_seed(1234)
    ref.densify_and_prune(thr, min_op, extent, max_screen_size)
    gen = torch.Generator(device=DEV).manual_seed(1234)
    stats = g.densify_and_prune(thr, min_op, extent, max_screen_size, generator=gen)
    print(stats, flush=True)
    assert stats["cloned"] > 100 and stats["split"] > 100 and stats["total"] < 40_000 + stats["cloned"] + stats["split"]
    _compare(g, ref)


def test_prune_and_densify_alone_and_a_step_after():
    from gs_hip.loss import fused_photometric_loss
    from gs_hip.model import render
    g, _ = _trained(seed=5)
    ref = RefModel(g)
    min_op = float(torch.quantile(torch.sigmoid(g._opacity.detach()), 0.1))
    ref.prune(min_op, 5.0, 30)
    g.prune(0.0, min_op, 5.0, 30)
    _compare(g, ref)
    # densify only: no prune pass
    grads = (g.xyz_gradient_accum / g.denom).nan_to_num(0.0)
    thr = float(torch.quantile(grads[grads > 0], 0.8))
    extent = float(torch.exp(g._scaling.detach()).max(dim=1).values.median()) / 0.01
    torch.cuda.manual_seed(77)
    r = ref.accum / ref.denom
    r[r.isnan()] = 0.0
    ref.clone(r, thr, extent)
    ref.split(r, thr, extent)
    g.densify(thr, 0.0, extent, generator=torch.Generator(device=DEV).manual_seed(77))
    _compare(g, ref)
    # the new set trains: a step with the Adam step inside the backward
    sc = scenes.training_scene(seed=5, P=40_000, W=320, H=180)
    out = render(g, sc["cam"].to(DEV), sc["bg"].to(DEV), clamp=False, adam_in_backward=True)
    loss, _ = fused_photometric_loss(out["image"], sc["target"].to(DEV), clamp01=True)
    loss.backward()
    g.optimizer.step()
    g.add_densification_stats(out["viewspace_points"], out["radii"])
    torch.cuda.synchronize()
    assert torch.isfinite(g._xyz).all() and int((g.denom > 0).sum()) > 1000


def test_reset_opacity():
    g, _ = _trained(seed=6, steps=2)
    before = g._opacity.detach().clone()
    step = g.optimizer.state[g._opacity]["step"]
    g.reset_opacity()
    x = torch.min(torch.sigmoid(before), torch.ones_like(before) * 0.01)
    assert torch.equal(g._opacity.detach(), torch.log(x / (1 - x)))
    st = g.optimizer.state[g._opacity]
    assert st["step"] == step and not st["exp_avg"].any() and not st["exp_avg_sq"].any()
    assert any(g._opacity is p for grp in g.optimizer.param_groups for p in grp["params"])


def test_ply_round_trip_on_the_device(tmp_path):
    g, _ = _trained(seed=7, steps=1)
    path = str(tmp_path / "point_cloud.ply")
    g.save_ply(path)
    h = GaussianParams.from_ply(path, device=DEV)
    for n in ("_xyz", "_features", "_opacity", "_scaling", "_rotation"):
        assert torch.equal(getattr(g, n).detach(), getattr(h, n).detach()), n
    assert h.active_sh_degree == 3
