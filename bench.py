"""Benchmark: vertex-substeps/s of the XPBD cloth step on one MI355X.

Workload (default): a 1024 x 1024 vertex cloth (1,048,576 vertices, 2,093,058
triangles, 3,141,633 stretch and 3,137,541 bending constraints), FP32, one frame
= 20 substeps x 2 Jacobi iterations at dt = 1/60 s.  A "step" is one frame.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line (rank 0).  Inputs are resident in HBM before timing starts.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=1024, help="cloth side (vertices)")
    p.add_argument("--substeps", type=int, default=20)
    p.add_argument("--iters", type=int, default=2)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    return p.parse_args()


def build_scene(slx, n):
    import torch
    g = slx.scene.grid_cloth(n, size=1.0, height=0.5, jitter=1e-3 / max(1, n // 64), seed=0)
    mesh = slx.Mesh(x=g.x, faces=g.faces, device="cuda")
    sim = slx.Solver(mesh, None, dHat=2.5e-5, stiffness_stretch=1e6, stiffness_bending=1e3,
                     g=(0.0, -9.8, 0.0), dt=1.0 / 60.0)
    sel = torch.zeros(mesh.num_verts, device="cuda")
    sel[0] = 1
    sel[(n - 1) * n] = 1
    sim.set_fixed_vertices(sel)
    return mesh, sim


def cpu_baseline(mesh, substeps, iters, seconds):
    """FP32 serial oracle (1 core) on the same cloth: whole substeps
    (predict, `iters` spring passes, velocity/position update), repeated for
    about `seconds`."""
    from oracle import oracle as orc
    d = mesh.host_dict()
    ref = orc.OracleCloth(d, None, dhat=2.5e-5, stiffness_stretch=1e6, stiffness_bending=1e3,
                          g=(0.0, -9.8, 0.0), dt=1.0 / 60.0, dtype=np.float32)
    ref.dt = (1.0 / 60.0) / substeps     # forward(1) then advances one substep of dt/substeps
    t0 = time.perf_counter()
    reps = 0
    while True:
        ref.forward(1, iters)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return dict(value=mesh.num_verts * reps / el, unit="vertex-substeps/s", cores=1, kind="port",
                sample=f"{reps} substeps x {iters} Jacobi iterations of the {mesh.num_verts}-vertex cloth "
                       f"on the serial FP32 oracle ({el:.1f} s)")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)
    import torch
    slx = importlib.import_module("starlab-physics_amd")
    from importlib import import_module
    rl = import_module("starlab-physics_amd.roofline")
    abi = import_module("starlab-physics_amd._abi")

    mesh, sim = build_scene(slx, a.n)
    V, E, B, F = mesh.num_verts, mesh.num_edges, mesh.bending_constraint_count, mesh.num_faces
    for _ in range(a.warmup):
        sim.forward(a.substeps, a.iters)
    torch.cuda.synchronize()

    launches_per_frame = a.substeps * (a.iters + 2) + 4
    abi.timer_enable(launches_per_frame * a.steps + 64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.forward(a.substeps, a.iters)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern = {}
    for name in abi.timer_names():
        tot, cnt = abi.timer_read(name)
        kern[name] = (tot, cnt)
    abi.timer_enable(0)

    value = V * a.substeps * a.steps / el
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, cnt = kern[dom]
    avg_s = tot_ms / cnt / 1e3
    nbytes = rl.kernel_bytes(dom, V, E, B)
    achieved = nbytes / avg_s / 1e9
    out = {
        "metric": "vertex-substeps/s",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded grid cloth)",
        "config": {"workload": f"{a.n}x{a.n} cloth ({V} vertices, {F} triangles, {E} stretch + {B} "
                               f"bending constraints), {a.substeps} substeps x {a.iters} Jacobi "
                               f"iterations per frame, collision off",
                   "substeps_per_step": a.substeps, "iters": a.iters, "collision": False},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None, "kernel": dom,
                     "algorithmic_bytes_per_launch": nbytes, "avg_launch_us": avg_s * 1e6},
        "kernels_ms": {k: {"total_ms": v[0], "launches": v[1], "avg_us": v[0] / max(v[1], 1) * 1e3}
                       for k, v in kern.items()},
        "step_roofline_frac": (rl.substep_bytes(V, E, B, n_iters=a.iters) * a.substeps * a.steps / el)
                              / rl.HBM_PEAK_BYTES_PER_S,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, a.substeps, a.iters, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
