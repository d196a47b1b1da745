"""Algorithmic HBM byte model of the XPBD cloth step (FP32 / int32), the single
source of truth for roofline figures.

Per vertex-array pass each array is counted once; on-chip temporaries (dx, nc)
are not counted.  V/E/B/F are the dynamic mesh's vertex, edge, bending and face
counts; c is the mean number of broad-phase candidates per vertex.

  predict                      40 V   (read x, v, f; write y)
  one Jacobi spring pass       12 (E+B) + 40 V   (ids + rest length; read y, x, f*w; write y)
  velocity / position update   52 V   (read x, y, f; write v, x)
  refit                        76 F + 12 V
  traversal                    16 V + 4 c V + 32 (2F - 1)
  narrow phase                 16 c V + 24 V
  per-frame rebuild            Morton 20 F + 12 V, radix sort 68 F, Karras 16 F
"""
from __future__ import annotations

HBM_PEAK_BYTES_PER_S = 8.0e12      # MI355X HBM3E


def predict(V):
    return 40 * V


def spring_pass(V, E, B):
    return 12 * (E + B) + 40 * V


def update_vx(V):
    return 52 * V


def refit(V, F):
    return 76 * F + 12 * V


def traversal(V, F, c):
    return 16 * V + 4 * c * V + 32 * (2 * F - 1)


def narrow(V, c):
    return 16 * c * V + 24 * V


def rebuild(V, F):
    return (20 * F + 12 * V) + 68 * F + 16 * F


def kernel_bytes(name: str, V, E, B) -> int:
    """Algorithmic bytes of one launch of a Jacobi kernel variant."""
    sp = spring_pass(V, E, B)
    return {
        "compute_y": predict(V),
        "spring": sp,
        "spring_pred": predict(V) + sp,
        "spring_final": sp + update_vx(V),
        "spring_pred_final": predict(V) + sp + update_vx(V),
        "update_vx": update_vx(V),
    }[name]


def substep_bytes(V, E, B, F=0, n_iters=1, collision=False, c=0.0, substeps_per_frame=20):
    """Bytes of one substep; with collision, the per-frame rebuild is amortised
    over the frame's substeps."""
    b = predict(V) + n_iters * spring_pass(V, E, B) + update_vx(V)
    if collision:
        b += refit(V, F) + n_iters * (traversal(V, F, c) + narrow(V, c))
        b += rebuild(V, F) / substeps_per_frame
    return b
