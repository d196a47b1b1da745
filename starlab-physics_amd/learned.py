"""Learned warm start for the Jacobi solve.

A per-vertex MLP (24 -> 64 -> 64 -> 64 -> 3, GELU; BF16 weights and
activations on MFMA, FP32 accumulation; csrc/mlp.hip) predicts the correction
y* - y_pred that plain Jacobi iterations would converge to within a substep,
from 1-ring strain and velocity features.  The prediction seeds the iterate;
plain Jacobi passes (XPBD.py:140-184, 383-387) continue until the update
residual max_v |y_{k+1} - y_k|_inf is <= tol; if the residual rises between two
iterations the substep restarts from y_pred without the network.  The training
data come from the plain solver itself, run to a tighter residual, and the
network is trained on the GPU by the library's own forward/backward kernels
(slx_mlp_grad) with Adam (slx_mlp_adam).
"""
from __future__ import annotations

import time

import numpy as np

from . import _abi

N_FEAT = 24
N_PARAMS = None


def n_params():
    global N_PARAMS
    if N_PARAMS is None:
        N_PARAMS = _abi.lib().slx_mlp_num_params()
    return N_PARAMS


# parameter layout (csrc/mlp.hip): W1[64][32] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[32][64] b4[32]
LAYOUT = [("W1", (64, 32)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (32, 64)), ("b4", (32,))]
REAL = {"W1": (64, 24), "W4": (3, 64), "b4": (3,)}


def unpack(flat):
    out, o = {}, 0
    for name, shp in LAYOUT:
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp)
        o += n
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    flat = np.zeros(n_params(), np.float32)
    P = unpack(flat)
    for name, fan_in, shape in (("W1", 24, (64, 24)), ("W2", 64, (64, 64)), ("W3", 64, (64, 64)), ("W4", 64, (3, 64))):
        P[name][:shape[0], :shape[1]] = rng.normal(0, np.sqrt(2.0 / fan_in), shape).astype(np.float32)
    P["W4"][:3] *= 0.1
    return flat


class WarmStart:
    """Model state on the device: FP32 master weights, BF16 working copy, Adam
    moments, feature normalisation, output scale."""

    def __init__(self, device="cuda", seed=0):
        import torch
        self.device = device
        self.w = torch.from_numpy(init_params(seed)).to(device)
        self.wb = torch.zeros(n_params(), dtype=torch.bfloat16, device=device)
        self.m1 = torch.zeros_like(self.w)
        self.m2 = torch.zeros_like(self.w)
        self.fmean = torch.zeros(N_FEAT, dtype=torch.float32, device=device)
        self.fistd = torch.ones(N_FEAT, dtype=torch.float32, device=device)
        self.out_scale = 1.0
        self.step = 0
        self.sync_bf16()

    def sync_bf16(self):
        _abi.call("slx_mlp_to_bf16", _abi.ptr(self.w), _abi.ptr(self.wb), _abi.stream_handle())

    def set_normalisation(self, feat, target):
        import torch
        self.fmean.copy_(feat.mean(dim=0))
        self.fistd.copy_(1.0 / (feat.std(dim=0) + 1e-6))
        self.out_scale = float(target.abs().mean().item() * 3.0 + 1e-12)

    def grad(self, feat, target_scaled, n_blocks=None):
        """Gradient of the mean squared error (over samples, 3 outputs) w.r.t.
        the FP32 parameters, by the library's MFMA forward/backward kernel."""
        import torch
        n = feat.shape[0]
        tiles = (n + 31) // 32
        nb = n_blocks or max(1, min(256, tiles))
        part = torch.empty((nb, n_params()), dtype=torch.float32, device=self.device)
        lp = torch.empty(nb, dtype=torch.float32, device=self.device)
        g = torch.empty(n_params(), dtype=torch.float32, device=self.device)
        _abi.call("slx_mlp_grad", n, _abi.ptr(feat), _abi.ptr(target_scaled), _abi.ptr(self.wb), _abi.ptr(self.w),
                  _abi.ptr(self.fmean), _abi.ptr(self.fistd), nb, _abi.ptr(part), _abi.ptr(lp), _abi.ptr(g),
                  _abi.stream_handle())
        return g, lp.sum()

    def train(self, feat, target, epochs=40, batch=8192, lr=2e-3, seed=0, log=None):
        """Minibatch Adam on the GPU; target in metres (scaled by out_scale here)."""
        import torch
        n = feat.shape[0]
        ts = (target / self.out_scale).contiguous()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            tot = 0.0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                f = feat[idx].contiguous()
                t = ts[idx].contiguous()
                tiles = (batch + 31) // 32
                nb = min(256, tiles)
                part = torch.empty((nb, n_params()), dtype=torch.float32, device=self.device)
                lp = torch.empty(nb, dtype=torch.float32, device=self.device)
                _abi.call("slx_mlp_grad", batch, _abi.ptr(f), _abi.ptr(t), _abi.ptr(self.wb), _abi.ptr(self.w),
                          _abi.ptr(self.fmean), _abi.ptr(self.fistd), nb, _abi.ptr(part), _abi.ptr(lp), None,
                          _abi.stream_handle())
                self.step += 1
                _abi.call("slx_mlp_adam", nb, _abi.ptr(part), _abi.ptr(self.w), _abi.ptr(self.wb), _abi.ptr(self.m1),
                          _abi.ptr(self.m2), float(lr), 0.9, 0.999, 1e-8, self.step, _abi.stream_handle())
                tot += float(lp.sum().item())
            losses.append(tot / max(1, n // batch))
            if log:
                log(ep, losses[-1])
        return losses


# ----------------------------------------------------------------- solving

class ResidualSolver:
    """Per-substep Jacobi to a residual, plain or warm-started, on a Solver's
    mesh (collision off)."""

    def __init__(self, sim, model: WarmStart | None = None):
        import torch
        self.sim = sim
        self.m = sim.mesh_dy
        self.model = model
        nv = self.m.num_verts
        self.ya = self.m.y4
        self.yb = torch.empty_like(self.m.y4)
        self.ypred = torch.empty_like(self.m.y4)
        self.res = torch.zeros(1, dtype=torch.int32, device=self.m.x4.device)
        self.inv_s = float(1.0 / np.mean(self.m.l0_np))
        self.feat = torch.empty((nv, N_FEAT), dtype=torch.float32, device=self.m.x4.device)

    def _spring(self, a, b, s):
        m = self.m
        _abi.call("slx_spring_pass", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.fixed), _abi.ptr(a), _abi.ptr(b),
                  _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width), _abi.ptr(m.adj_entries), s["k_s"],
                  s["k_b"], _abi.stream_handle())

    def _residual(self, a, b):
        self.res.zero_()
        _abi.call("slx_max_update", self.m.num_verts, _abi.ptr(a), _abi.ptr(b), _abi.ptr(self.res),
                  _abi.stream_handle())
        return float(self.res.cpu().numpy().view(np.float32)[0])

    def predict(self, s):
        m = self.m
        _abi.call("slx_compute_y", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.ypred), s["dt"],
                  *s["g"], _abi.stream_handle())

    def features(self, s):
        m = self.m
        _abi.call("slx_features", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.ypred),
                  _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width), _abi.ptr(m.adj_entries), s["dt"],
                  self.inv_s, _abi.ptr(self.feat), _abi.stream_handle())
        return self.feat

    def warm(self, s):
        m, md = self.m, self.model
        _abi.call("slx_warm_start", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.ypred),
                  _abi.ptr(self.ya), _abi.ptr(m.fixed),
                  _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width), _abi.ptr(m.adj_entries), s["dt"],
                  self.inv_s, _abi.ptr(md.wb), _abi.ptr(md.w), _abi.ptr(md.fmean), _abi.ptr(md.fistd),
                  float(md.out_scale), _abi.stream_handle())

    def iterate(self, s, tol, max_iters, allow_fallback):
        """Jacobi from ya until the update residual <= tol; returns (iters, ok, fell_back)."""
        prev = None
        for k in range(1, max_iters + 1):
            self._spring(self.ya, self.yb, s)
            r = self._residual(self.ya, self.yb)
            self.ya, self.yb = self.yb, self.ya
            if r <= tol:
                return k, True, False
            if allow_fallback and (not np.isfinite(r) or (prev is not None and r > prev)):
                return k, False, True
            prev = r
        return max_iters, False, False

    def substep(self, n_substeps, tol=1e-5, max_iters=2000, learned=False):
        """One substep solved to the residual; returns (iterations, fell_back)."""
        s = self.sim.scalars(n_substeps)
        self.predict(s)
        self.ya.copy_(self.ypred)
        fell = False
        iters = 0
        if learned and self.model is not None:
            self.warm(s)
            k, ok, fell = self.iterate(s, tol, max_iters, allow_fallback=True)
            iters += k
            if not ok:
                self.ya.copy_(self.ypred)
                fell = True
        if not learned or fell:
            k, _, _ = self.iterate(s, tol, max_iters, allow_fallback=False)
            iters += k
        m = self.m
        self._finish(s)
        return iters, fell

    def _finish(self, s):
        """velocity/position update from the solved y; keep mesh.y4 = y."""
        m = self.m
        _abi.call("slx_update_vx", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.ya), _abi.ptr(m.fixed),
                  s["damping"], s["dt"], _abi.stream_handle())
        if self.ya.data_ptr() != m.y4.data_ptr():
            m.y4.copy_(self.ya)
            self.yb = self.ya
            self.ya = m.y4

    def collect(self, n_substeps, tol, max_iters, n_samples, rng, out_f, out_t):
        """One substep of plain Jacobi to `tol`, recording (features, y* - y_pred)
        for n_samples random free vertices."""
        import torch
        s = self.sim.scalars(n_substeps)
        self.predict(s)
        self.features(s)
        self.ya.copy_(self.ypred)
        k, ok, _ = self.iterate(s, tol, max_iters, allow_fallback=False)
        free = torch.nonzero(self.m.fixed > 0).flatten()
        pick = free[torch.from_numpy(rng.choice(len(free), min(n_samples, len(free)), replace=False)).to(free.device)]
        out_f.append(self.feat[pick].clone())
        out_t.append((self.ya[pick, :3] - self.ypred[pick, :3]).clone())
        self._finish(s)
        return k, ok
