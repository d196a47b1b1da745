"""GPU: a frame captured into a HIP graph (Solver.capture) and replayed equals the
same frames run eagerly, bit for bit, replay after replay -- with eager work and
host synchronisations between the replays: the reference's own case (64x64
cloth on an icosphere, cell LBVH rebuilt per frame) and the benchmark's drape at
64x64 (triangle tree refit per substep, candidate cache, velocity pass, three
streams)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sphere(slx):
    import torch
    n = 64
    g = slx.scene.grid_cloth(n, size=1.0, height=0.5, jitter=1e-3, seed=0)
    cloth = slx.Mesh(x=g.x, faces=g.faces)
    ico = slx.scene.icosphere(4, 0.3)
    obst = slx.Mesh(x=ico.x, faces=ico.faces, is_static=True)
    sim = slx.Solver(cloth, obst, dHat=1e-4, stiffness_stretch=1e6, stiffness_bending=1e3, g=(0.0, -9.8, 0.0),
                     dt=1.0 / 60.0)
    sim.enable_collision_handling = True
    sim.bvh_mode = "cell"
    sim.padding = 0.05
    sel = torch.zeros(cloth.num_verts, device="cuda")
    sel[0] = 1
    sel[(n - 1) * n] = 1
    sim.set_fixed_vertices(sel)
    return cloth, sim, (20, 1)


def drape(slx):
    import bench
    cloth, terr, sim = bench.build_scene(slx, 64, True)
    sim.enable_velocity_update = True
    return cloth, sim, (20, 2)


@pytest.mark.parametrize("scene", ["sphere", "drape"])
def test_graph_replay_equals_eager(slx, scene):
    import torch
    make = sphere if scene == "sphere" else drape
    ca, sa, args = make(slx)
    cb, sb, _ = make(slx)
    g = sb.capture(*args)                      # (runs one eager frame first)
    sa.forward(*args)
    for f in range(6):
        print(f"[graph] scene={scene} frame={f}", flush=True)
        g.replay()
        sa.forward(*args)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(cb.positions(), ca.positions(), err_msg=f"x, frame {f}")
        np.testing.assert_array_equal(cb.v4[:, :3].cpu().numpy(), ca.v4[:, :3].cpu().numpy(), err_msg=f"v, frame {f}")
        assert sb.status() == sa.status()
    assert not np.allclose(cb.positions(), make(slx)[0].positions())     # the frames moved the cloth
