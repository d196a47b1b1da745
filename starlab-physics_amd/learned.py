"""Learned warm start for the Jacobi solve.

A per-vertex MLP (24 -> 64 -> 64 -> 64 -> 3, GELU; BF16 weights and
activations on MFMA, FP32 accumulation; csrc/mlp.hip) predicts the correction
y* - y1 that the substep's residual solve converges to (y1: the result of the
reference Jacobi pass, XPBD.py:140-184, 383-387; y*: the solution of the
implicit-Euler stretch problem anchored at y1, csrc/solve.hip), from 1-ring
strain and velocity features.  The prediction seeds the iterate; plain Jacobi
passes continue until the residual (largest Jacobi step) is <= tol, and if the
residual rises the substep restarts from y1 without the network.  The training
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
        """Per-feature mean and 1/std of the training set (features with no
        spread get weight 0, not 1/eps); out_scale = 3 x mean |target|."""
        import torch
        mean = feat.mean(dim=0)
        std = feat.std(dim=0)
        scale = mean.abs().clamp_min(1.0)
        self.fmean.copy_(mean)
        self.fistd.copy_(torch.where(std > 1e-6 * scale, 1.0 / std.clamp_min(1e-30), torch.zeros_like(std)))
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
    """Per-substep solve to a residual, plain or warm-started, on a Solver's mesh
    (collision off).

    A substep: predict and one reference Jacobi pass (XPBD.py:443-469 with one
    iteration) give y1; then Jacobi passes on the substep's implicit-Euler stretch
    problem anchored at y1 (slx_implicit_passes) run until the residual (largest
    Jacobi step) is <= tol.  Warm-started, the network's correction is added to y1
    before the passes.  Passes are launched in chunks with no host
    synchronisation inside a chunk (they stop by themselves once converged); the
    residual of every pass comes back once per chunk.  If the residual at the end
    of a chunk is not finite or larger than at the end of the previous chunk, the
    warm-started solve is abandoned and the substep restarts from y1 without the
    network (the plain solver is the fallback and the judge of the residual)."""

    def __init__(self, sim, model: WarmStart | None = None, chunk=4):
        import torch
        self.sim = sim
        self.m = sim.mesh_dy
        self.model = model
        self.chunk = int(chunk)
        nv = self.m.num_verts
        dev = self.m.x4.device
        self.y1 = torch.empty_like(self.m.y4)
        self.ya = self.m.y4
        self.yb = torch.empty_like(self.m.y4)
        self.ypred = torch.empty_like(self.m.y4)
        self.res = torch.zeros(1, dtype=torch.int32, device=dev)
        self.conv = torch.zeros(1, dtype=torch.int32, device=dev)
        self.inv_s = float(1.0 / np.mean(self.m.l0_np))
        self.feat = torch.empty((nv, N_FEAT), dtype=torch.float32, device=dev)
        self.kappa = None
        # running mean of the passes each mode needed: the first chunk of a solve is
        # sized from it, so a converging solve usually synchronises once
        self._expect = {False: None, True: None}

    def _scalars(self, n_substeps):
        s = self.sim.scalars(n_substeps)
        # the stretch constraints' stiffness * dt^2: the reference passes it as k_s
        # (XPBD.py:354-355)
        self.kappa = s["k_s"]
        return s

    def reference_pass(self, s):
        """y1 = one reference Jacobi pass (bending on x, stretch on y) from the prediction."""
        m = self.m
        _abi.call("slx_compute_y", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.ypred), s["dt"],
                  *s["g"], _abi.stream_handle())
        _abi.call("slx_spring_pass", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.fixed), _abi.ptr(self.ypred),
                  _abi.ptr(self.y1), _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width),
                  _abi.ptr(m.adj_entries), s["k_s"], s["k_b"], _abi.stream_handle())

    def features(self, s):
        m = self.m
        _abi.call("slx_features", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.y1),
                  _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width), _abi.ptr(m.adj_entries), s["dt"],
                  self.inv_s, _abi.ptr(self.feat), _abi.stream_handle())
        return self.feat

    def warm(self, s):
        """ya = y1 + f * network(features(y1)) * out_scale (one fused kernel)."""
        m, md = self.m, self.model
        _abi.call("slx_warm_start", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(self.y1),
                  _abi.ptr(self.ya), _abi.ptr(m.fixed),
                  _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width), _abi.ptr(m.adj_entries), s["dt"],
                  self.inv_s, _abi.ptr(md.wb), _abi.ptr(md.w), _abi.ptr(md.fmean), _abi.ptr(md.fistd),
                  float(md.out_scale), _abi.stream_handle())

    def solve(self, tol, max_iters, allow_fallback, mode=False):
        """Jacobi passes from ya to residual <= tol.  Returns (iterations, converged,
        gave_up, result buffer, residual).  The first chunk covers the passes this
        mode needed recently (+2, one of them measures the answer); later chunks
        double from `chunk`."""
        import torch
        m = self.m
        if self.res.numel() < max_iters + 1:
            self.res = torch.zeros(max_iters + 1, dtype=torch.int32, device=m.x4.device)
        _abi.call("slx_solve_reset", max_iters + 1, _abi.ptr(self.res), _abi.ptr(self.conv), _abi.stream_handle())
        k0, prev = 0, None
        e = self._expect.get(mode)
        nxt = int(np.ceil(e)) + 2 if e is not None else self.chunk
        grow = self.chunk
        while k0 < max_iters:
            cnt = min(nxt, max_iters - k0)
            nxt = grow
            grow *= 2
            _abi.call("slx_implicit_passes", m.num_verts, _abi.ptr(m.fixed), _abi.ptr(self.y1), _abi.ptr(self.ya),
                      _abi.ptr(self.yb), _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width),
                      _abi.ptr(m.adj_entries), self.kappa, k0, cnt, float(tol), _abi.ptr(self.res),
                      _abi.ptr(self.conv), _abi.stream_handle())
            r = self.res[k0:k0 + cnt].cpu().numpy().view(np.float32)
            hit = np.nonzero(r <= tol)[0]
            if len(hit):
                c = k0 + int(hit[0])
                self._expect[mode] = c if e is None else 0.7 * e + 0.3 * c
                return c, True, False, (self.yb if c & 1 else self.ya), float(r[hit[0]])
            last = float(r[-1])
            if allow_fallback and (not np.isfinite(last) or (prev is not None and last > prev)):
                return k0 + cnt, False, True, None, last
            prev = last
            k0 += cnt
        return max_iters, False, False, (self.yb if max_iters & 1 else self.ya), prev

    def substep(self, n_substeps, tol=1e-5, max_iters=2000, learned=False):
        """One substep solved to the residual; returns (iterations, fell_back).
        Iterations count the Jacobi passes of the residual solve (the reference
        pass and the passes spent before a fallback included)."""
        s = self._scalars(n_substeps)
        self.reference_pass(s)
        fell = False
        iters = 0
        y = None
        if learned and self.model is not None:
            self.warm(s)
            k, ok, gave_up, y, _ = self.solve(tol, max_iters, allow_fallback=True, mode=True)
            iters += k
            fell = gave_up
        if not learned or self.model is None or fell:
            self.ya.copy_(self.y1)
            k, ok, _, y, _ = self.solve(tol, max_iters, allow_fallback=False)
            iters += k
        self._finish(s, y)
        return iters, fell

    def _finish(self, s, y):
        """velocity/position update from the solved y; mesh.y4 = y."""
        m = self.m
        _abi.call("slx_update_vx", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(y), _abi.ptr(m.fixed),
                  s["damping"], s["dt"], _abi.stream_handle())
        if y.data_ptr() != m.y4.data_ptr():
            m.y4.copy_(y)

    def collect(self, n_substeps, tol, max_iters, n_samples, rng, out_f, out_t):
        """One plain substep solved to `tol`, recording (features at y1, y* - y1)
        for n_samples random free vertices; returns (iterations, converged)."""
        import torch
        s = self._scalars(n_substeps)
        self.reference_pass(s)
        self.features(s)
        self.ya.copy_(self.y1)
        k, ok, _, y, _ = self.solve(tol, max_iters, allow_fallback=False)
        free = torch.nonzero(self.m.fixed > 0).flatten()
        pick = free[torch.from_numpy(rng.choice(len(free), min(n_samples, len(free)), replace=False)).to(free.device)]
        out_f.append(self.feat[pick].clone())
        out_t.append((y[pick, :3] - self.y1[pick, :3]).clone())
        self._finish(s, y)
        return k, ok
