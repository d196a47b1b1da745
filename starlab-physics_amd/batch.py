"""Batched cloth instances in one solver, one captured HIP graph per frame.

Many independent cloths of uneven sizes, each with its own obstacle, are packed
CSR-style into one dynamic and one static mesh (Mesh.from_instances; the
reference's only multi-mesh mechanism is the concatenation of
framework/meshio/concat.py:11-58).  Constraints never cross instances, so the
packed mesh runs through the ordinary frame (slx_frame); the broad phase keeps
one subtree per instance of a single LBVH built with one segmented sort
(slx_seg_build), so no instance sees another's triangles.

A frame is captured once into a HIP graph (torch.cuda.CUDAGraph) and replayed,
which removes every per-kernel launch from the host loop.  `autotune` picks the
spring kernels' launch shape (threads per block x vertices per thread) for this
scene's solver by timing replays of the captured frame.
"""
from __future__ import annotations

import time

import numpy as np

from . import scene as _scene
from .mesh import Mesh
from .solver import Solver


def instance_sides(n_inst=256, lo=16, hi=256, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, n_inst)


def make_instances(n_inst=256, side_lo=16, side_hi=256, seed=0, size=1.0, sphere_subdiv=3,
                   r_lo=0.2, r_hi=0.5, gap=0.05):
    """Cloth instance i: side_i x side_i vertex grid (side ~ U{side_lo..side_hi}),
    1 m square at height r_i + gap above its own sphere of radius r_i ~ U(r_lo, r_hi)
    at the origin.  Instances share the space but never interact."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(side_lo, side_hi + 1, n_inst)
    radii = rng.uniform(r_lo, r_hi, n_inst)
    base = _scene.icosphere(sphere_subdiv, 1.0)
    cloths, spheres = [], []
    for i in range(n_inst):
        r = float(radii[i])
        cloths.append(_scene.grid_cloth(int(sides[i]), size=size, height=r + gap, jitter=1e-4, seed=seed + i))
        spheres.append(_scene.TriMesh((base.x.astype(np.float64) * r).astype(np.float32), base.faces))
    return cloths, spheres, sides, radii


class BatchedScene:
    def __init__(self, cloths, spheres, *, dt=1.0 / 60.0, stiffness_stretch=1e6, stiffness_bending=1e3,
                 thickness=5e-3, padding=None, g=(0.0, -9.8, 0.0), device="cuda"):
        import torch
        self.cloth = Mesh.from_instances(cloths, device=device)
        self.obst = Mesh.from_instances(spheres, is_static=True, device=device)
        self.n_inst = len(cloths)
        sim = Solver(self.cloth, self.obst, dHat=thickness * thickness, stiffness_stretch=stiffness_stretch,
                     stiffness_bending=stiffness_bending, g=g, dt=dt)
        sim.enable_collision_handling = True
        sim.bvh_mode = "tri"
        sim.refit_every_substep = True
        sim.padding = 0.5 * thickness if padding is None else padding
        # pin two corners of every instance
        sel = torch.zeros(self.cloth.num_verts, device=device)
        vb = self.cloth.inst_np["vert_base"]
        for i, c in enumerate(cloths):
            n = int(round(np.sqrt(len(c.x))))
            sel[int(vb[i])] = 1
            sel[int(vb[i]) + n - 1] = 1
        sim.set_fixed_vertices(sel)
        # the cloths fall onto their spheres: the obstacle set's cache is rebuilt
        # every substep unless the autotuner finds otherwise
        sim.static_cache_every = 1
        self.sim = sim
        self.graph = None
        self.shape = (256, 1)

    # ------------------------------------------------------------------ frames
    def forward(self, n_substeps, n_iters):
        self.sim.forward(n_substeps, n_iters)

    def capture(self, n_substeps, n_iters):
        """Capture one frame into a HIP graph (buffers are allocated and the static
        tree built by an eager warm-up frame first)."""
        g = self.sim.capture(n_substeps, n_iters)      # (an eager frame first: allocations, static tree)
        self.graph = g
        self.graph_shape = (n_substeps, n_iters)
        return g

    def replay(self):
        self.graph.replay()

    # --------------------------------------------------------------- autotune
    def _time_frame(self, state, n_substeps, n_iters, reps):
        import torch
        self.cloth.x4.copy_(state[0])
        self.cloth.v4.copy_(state[1])
        self.capture(n_substeps, n_iters)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def autotune(self, n_substeps, n_iters, blocks=(64, 128, 256), vpts=(1, 2, 4), reps=3, knobs=True):
        """Time `reps` graph replays of one frame for every launch shape of the
        spring kernels (threads per block x vertices per thread) and keep the
        fastest; then, one knob at a time with the others at their best, the
        first-tier walk's stack depth, the candidate cache's margin (a multiple
        of the padding), the refit's subtree size and the interval of the
        obstacle set's cache.  Each setting is captured
        afresh (launch shapes are baked into the graph).  No knob changes the
        results.  Returns (best spring shape, {shape: seconds per frame}); the
        other knobs' timings are in self.tuning."""
        results = {}
        state = [t.clone() for t in (self.cloth.x4, self.cloth.v4)]
        for bs in blocks:
            for vpt in vpts:
                self.sim.spring_launch = (bs, vpt)
                results[(bs, vpt)] = self._time_frame(state, n_substeps, n_iters, reps)
        best = min(results, key=results.get)
        self.sim.spring_launch = best
        self.tuning = {"spring_launch": {f"{k[0]}x{k[1]}": v for k, v in results.items()}}
        if knobs:
            pad = self.sim.padding
            for name, values, apply in (
                    ("walk_stack", (16, 24, 32, 64), lambda v: setattr(self.sim, "walk_stack", v)),
                    ("cache_margin", (0.125, 0.25, 0.5), lambda v: setattr(self.sim, "cache_margin", v * pad)),
                    ("refit_sub", (1024, 2048), lambda v: setattr(self.sim, "refit_shape", (v, 2048))),
                    ("static_cache_every", (1, 4), lambda v: setattr(self.sim, "static_cache_every", v))):
                t = {}
                for v in values:
                    apply(v)
                    t[v] = self._time_frame(state, n_substeps, n_iters, reps)
                b = min(t, key=t.get)
                apply(b)
                self.tuning[name] = {"best": b, "seconds": {str(k): x for k, x in t.items()}}
        self.cloth.x4.copy_(state[0])
        self.cloth.v4.copy_(state[1])
        self.capture(n_substeps, n_iters)
        self.shape = best
        return best, results
