"""Solver: the reference's XPBD.Solver surface over the gfx950 kernels.

Mirrors framework/physics/XPBD.py:9-106 (constructor and attributes),
forward (443-469), reset (108-112) and set_fixed_vertices (389-395).
Extensions: forward(n_substeps, n_iters) runs n_iters spring passes per
substep (the reference runs exactly one, XPBD.py:463); `step` is an alias.
"""
from __future__ import annotations

import numpy as np

from . import _abi


def _f32(v) -> float:
    return float(np.float32(v))


class Solver:
    def __init__(self, mesh_dy, mesh_st, dHat, stiffness_stretch, stiffness_bending, g, dt):
        import torch
        if not torch.cuda.is_available():
            raise _abi.SlxError("Solver needs a HIP device (torch.cuda.is_available() is False)")
        _abi.lib()
        self.mesh_dy = mesh_dy
        self.mesh_st = mesh_st
        self.g = tuple(float(c) for c in g)
        self.dt = dt
        self.dHat = dHat
        self.stiffness_stretch = stiffness_stretch
        self.stiffness_bending = stiffness_bending
        self.damping = 0.001
        self.mu = 0.8
        self.padding = 0.05
        self.enable_velocity_update = False
        self.enable_collision_handling = False
        self.enable_move_obstacle = False
        self.export_mesh = False
        self.max_num_verts_dy = mesh_dy.num_verts
        self.max_num_edges_dy = mesh_dy.num_edges
        self.max_num_faces_dy = mesh_dy.num_faces
        self.max_num_verts_st = mesh_st.num_verts if mesh_st is not None else 0
        self.max_num_edges_st = mesh_st.num_edges if mesh_st is not None else 0
        self.max_num_faces_st = mesh_st.num_faces if mesh_st is not None else 0
        # build extensions
        self.n_iters = 1
        dev = mesh_dy.x4.device
        nv = mesh_dy.num_verts
        self._x_alt = torch.empty((nv, 4), dtype=torch.float32, device=dev)
        self._v_alt = torch.empty((nv, 4), dtype=torch.float32, device=dev)
        self._y_alt = torch.empty((nv, 4), dtype=torch.float32, device=dev)

    # ------------------------------------------------------------ reference API
    def reset(self):
        self.mesh_dy.reset()
        if self.mesh_st is not None:
            self.mesh_st.reset()

    def set_fixed_vertices(self, fixed_vertices):
        """XPBD.py:389-395: fixed = 0 where selection >= 1, else 1."""
        import torch
        m = self.mesh_dy
        sel = torch.as_tensor(fixed_vertices, device=m.x4.device)
        m.fixed.copy_(torch.where(sel >= 1, 0.0, 1.0).to(torch.float32))
        _abi.call("slx_set_fixed", m.num_verts, _abi.ptr(m.fixed), _abi.ptr(m.m_inv), _abi.ptr(m.x4),
                  _abi.ptr(m.v4), _abi.ptr(m.y4), _abi.stream_handle())

    def scalars(self, n_substeps):
        """Per-substep scalars exactly as the reference forms them: dt_sub and
        stiffness*dt_sub^2 in double, rounded to float32 at the kernel call
        (XPBD.py:446, 354-355 — stiffness names are swapped there)."""
        dt_sub = self.dt / n_substeps
        return dict(dt=_f32(dt_sub), k_s=_f32(self.stiffness_bending * dt_sub * dt_sub),
                    k_b=_f32(self.stiffness_stretch * dt_sub * dt_sub), damping=_f32(self.damping),
                    g=tuple(_f32(c) for c in self.g))

    def forward(self, n_substeps, n_iters=None):
        n_iters = self.n_iters if n_iters is None else n_iters
        if self.enable_collision_handling:
            raise NotImplementedError("collision path not built yet")
        self._forward_plain(n_substeps, n_iters)

    step = forward

    def _forward_plain(self, n_substeps, n_iters):
        m = self.mesh_dy
        s = self.scalars(n_substeps)
        _abi.call("slx_substeps_plain", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4),
                  _abi.ptr(self._x_alt), _abi.ptr(self._v_alt), _abi.ptr(m.y4), _abi.ptr(self._y_alt),
                  _abi.ptr(m.fixed), _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width),
                  _abi.ptr(m.adj_entries), n_substeps, n_iters, s["dt"], *s["g"], s["k_s"], s["k_b"],
                  s["damping"], _abi.stream_handle())

    # ----------------------------------------------------------- stage access
    def compute_y(self, dt):
        m = self.mesh_dy
        g = tuple(_f32(c) for c in self.g)
        _abi.call("slx_compute_y", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(m.y4),
                  _f32(dt), *g, _abi.stream_handle())

    def solve_spring_constraints(self, k_s, k_b):
        """One spring pass on y (in place through the alternate buffer)."""
        m = self.mesh_dy
        _abi.call("slx_spring_pass", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.fixed), _abi.ptr(m.y4),
                  _abi.ptr(self._y_alt), _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width),
                  _abi.ptr(m.adj_entries), _f32(k_s), _f32(k_b), _abi.stream_handle())
        m.y4.copy_(self._y_alt)

    def update_vx(self, dt):
        m = self.mesh_dy
        _abi.call("slx_update_vx", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(m.y4),
                  _abi.ptr(m.fixed), _f32(self.damping), _f32(dt), _abi.stream_handle())
