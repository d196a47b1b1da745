"""Float64 CPU reference of the learned warm start (features, MLP forward,
gradients).  TEST INFRASTRUCTURE ONLY: the checker of csrc/mlp.hip, imported by
tests; never by the product path.

The network's BF16 operands are emulated exactly (round-to-nearest-even of the
normalised input, of every weight matrix and of every hidden activation), so the
reference differs from the GPU only by FP32-vs-FP64 accumulation order, erf, and
the GPU backward's BF16 rounding of its own intermediate gradients.
"""
from __future__ import annotations

import numpy as np
import torch

CLIP = 8.0

LAYOUT = [("W1", (64, 32)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (32, 64)), ("b4", (32,))]


def features(x, v, y, adj, dt, inv_s):
    """24 features per vertex (csrc/mlp.hip header), float64, from the SELL rows
    of `adj` (topology.Adjacency) in row order."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    v = np.asarray(v, np.float64)
    nv = len(x)
    F = np.zeros((nv, 24))
    F[:, 0:3] = (y - x) * inv_s
    F[:, 3:6] = v * dt * inv_s
    off = adj.row_off
    for i in range(nv):
        ns = nb = 0
        for k in range(off[i], off[i + 1]):
            u, bend, l0 = int(adj.other[k]), bool(adj.kind[k]), float(adj.rest[k])
            d = (x[i] - x[u]) if bend else (y[i] - y[u])
            ln = np.linalg.norm(d)
            strain = (ln - l0) / l0
            n = d / ln if ln > 0 else np.zeros(3)
            if not bend:
                if ns < 6:
                    F[i, 6 + ns] = strain
                ns += 1
                F[i, 18:21] += strain * n
            else:
                if nb < 6:
                    F[i, 12 + nb] = strain
                nb += 1
                F[i, 21:24] += strain * n
    return F


def bf16(t):
    return t.to(torch.bfloat16).to(torch.float64)


def unpack(flat):
    flat = torch.as_tensor(np.asarray(flat, np.float32)).to(torch.float64)
    out, o = {}, 0
    for name, shp in LAYOUT:
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp).clone()
        o += n
    return out


def forward(P, feat, fmean, fistd, leaf_weights=False):
    """Output (N, 32) in float64 with BF16 operands emulated.  With
    leaf_weights, the BF16-rounded weights and the biases are leaf tensors with
    requires_grad (for gradients w.r.t. the values the GPU multiplies)."""
    f = torch.as_tensor(np.asarray(feat, np.float64))
    a = (f - torch.as_tensor(np.asarray(fmean, np.float64))) * torch.as_tensor(np.asarray(fistd, np.float64))
    a = a.clamp(-CLIP, CLIP)          # normalised inputs clamped to +-8 (csrc/mlp.hip norm_feature)
    a = torch.cat([a, torch.zeros(a.shape[0], 8, dtype=torch.float64)], dim=1)
    a = bf16(a)
    leaves = {}
    for name in ("W1", "W2", "W3", "W4"):
        w = bf16(P[name])
        if leaf_weights:
            w = w.clone().requires_grad_(True)
        leaves[name] = w
    for name in ("b1", "b2", "b3", "b4"):
        b = P[name].clone()
        if leaf_weights:
            b.requires_grad_(True)
        leaves[name] = b
    for l in (1, 2, 3):
        z = a @ leaves[f"W{l}"].T + leaves[f"b{l}"]
        a = bf16(torch.nn.functional.gelu(z))
    out = a @ leaves["W4"].T + leaves["b4"]
    return out, leaves


def loss_and_grad(flat, feat, target, fmean, fistd):
    """Mean over samples of the summed squared error on the 3 outputs and its
    gradient, in the flat parameter layout."""
    P = unpack(flat)
    out, leaves = forward(P, feat, fmean, fistd, leaf_weights=True)
    t = torch.as_tensor(np.asarray(target, np.float64))
    loss = ((out[:, :3] - t) ** 2).sum() / out.shape[0]
    loss.backward()
    g = np.concatenate([leaves[name].grad.numpy().reshape(-1) for name, _ in LAYOUT])
    return float(loss.item()), g
