"""Config-size GPU parity against the FP32 oracle, bit for bit:

* the self-collision stack: four 256 x 256 layers 0.02 m apart (262,144
  vertices, 520,200 triangles) resting on an icosphere (subdivision 4, 0.3 m),
  vertex-triangle self and obstacle collision at thickness 5 mm (dHat = 2.5e-5
  on squared distance), the triangle LBVH rebuilt every substep (a frame of one
  substep per call, dt = 1/600), ten substeps per 1/60 s, with Jacobi and with
  the graph-coloured Gauss-Seidel variant (XPBD.py:443-469,
  collision_constraints_x.py:5-310, graph_coloring.py:131-201);
* the batched form at its full size: 256 instances with sides drawn from
  16..256 (6,416,788 vertices) packed CSR-style (framework/meshio/concat.py:
  11-58), one frame of 20 substeps x 2 iterations, each of eight instances --
  the smallest, the largest and six more -- equal to the oracle simulating that
  instance alone."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def layers_scene(slx, method):
    import torch
    from oracle import oracle as orc
    n, layers = 256, 4
    g = slx.scene.stacked_layers(n, layers, 0.02, size=1.0, height=0.302, jitter=1e-3, seed=0)
    cloth = slx.Mesh(x=g.x, faces=g.faces)
    ico = slx.scene.icosphere(4, 0.3)
    obst = slx.Mesh(x=ico.x, faces=ico.faces, is_static=True)
    sim = slx.Solver(cloth, obst, dHat=0.005 ** 2, stiffness_stretch=1e6, stiffness_bending=1e3,
                     g=(0.0, -9.8, 0.0), dt=1.0 / 600.0)
    sim.method = method
    sim.enable_collision_handling = True
    sim.bvh_mode = "tri"
    sim.padding = 0.0025
    sim.records_per_vertex = 128        # a 5 mm thickness over a 3.9 mm grid: many contacts per vertex
    sel = torch.zeros(cloth.num_verts, device="cuda")
    for k in range(layers):
        sel[k * n * n] = 1
        sel[k * n * n + n - 1] = 1
    sim.set_fixed_vertices(sel)
    ref = orc.OracleCloth(cloth.host_dict(), dict(x=obst.positions(), faces=obst.fid_np, m_inv=obst.m_inv_np),
                          dhat=sim.dHat, stiffness_stretch=1e6, stiffness_bending=1e3, g=sim.g, dt=sim.dt,
                          padding=sim.padding, collision=True, bvh="tri", cap=sim.cache_size, method=method)
    return cloth, sim, ref


@pytest.mark.timeout(900)
@pytest.mark.parametrize("method", ["jacobi", "gs"])
def test_stacked_layers_256_bit_exact(slx, method):
    cloth, sim, ref = layers_scene(slx, method)
    if method == "gs":
        assert sim.coloring()["n_colors"] <= 16
    contacts = 0
    for k in range(20):                       # two frames of ten substeps, the tree rebuilt at each
        print(f"[layers] method={method} substep={k}", flush=True)
        before = ref.stats["overflow"]
        sim.clear_status()
        sim.forward(1, 1)
        ref.forward(1, 1)
        st = sim.status()
        assert st & ~1 == 0, (k, st)
        assert (st & 1 != 0) == (ref.stats["overflow"] > before), k
        c = sim.candidates()
        contacts += int((c["vt_st"][1] > 0).sum().item()) + int((c["vt_dy"][1] > 0).sum().item())
    np.testing.assert_array_equal(cloth.positions(), ref.x)
    np.testing.assert_array_equal(cloth.v4[:, :3].cpu().numpy(), ref.v)
    assert contacts > 0


@pytest.mark.timeout(900)
def test_batched_256_instances_one_frame(slx, orc):
    batch = importlib.import_module("starlab-physics_amd.batch")
    tb = importlib.import_module("test_gpu_batch")
    cloths, spheres, sides, radii = batch.make_instances(256, 16, 256, seed=0, gap=2e-3)
    bs = batch.BatchedScene(cloths, spheres, thickness=1e-3)
    assert bs.cloth.num_verts == 6416788
    vb = bs.cloth.inst_np["vert_base"]
    fixed = bs.cloth.fixed.cpu().numpy()
    rng = np.random.default_rng(0)
    pick = sorted({int(np.argmin(sides)), int(np.argmax(sides)), *rng.choice(256, 6, replace=False).tolist()})
    refs = {i: tb.oracle_instance(orc, cloths[i], spheres[i], bs, fixed[vb[i]:vb[i + 1]].copy()) for i in pick}
    print(f"[batched] instances {pick} sides {[int(sides[i]) for i in pick]}", flush=True)
    bs.capture(20, 2)                          # one eager frame, then the captured frame
    for r in refs.values():
        r.forward(20, 2)
    bs.replay()
    for r in refs.values():
        r.forward(20, 2)
    x = bs.cloth.positions()
    for i, r in refs.items():
        np.testing.assert_array_equal(x[vb[i]:vb[i + 1]], r.x, err_msg=f"instance {i} (side {sides[i]})")
    assert bs.sim.status() == 0
