"""Batched instances (CSR-packed cloths, one LBVH with a subtree per instance):
every instance evolves exactly as if it were simulated alone (bit-exact vs the
oracle per instance), graph replays equal eager frames, and the spring kernels'
launch shape does not change results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_batch(slx, n_inst=5, seed=3):
    batch = __import__("importlib").import_module("starlab-physics_amd.batch")
    cloths, spheres, sides, radii = batch.make_instances(n_inst, 6, 18, seed=seed, sphere_subdiv=2)
    return batch, cloths, spheres, batch.BatchedScene(cloths, spheres, thickness=0.01, padding=0.02)


def oracle_instance(orc, cloth_tm, sphere_tm, bs, fixed):
    import importlib
    slx = importlib.import_module("starlab-physics_amd")
    m = slx.Mesh(x=cloth_tm.x, faces=cloth_tm.faces, device=None)
    d = m.host_dict()
    d["fixed"] = fixed
    sim = bs.sim
    return orc.OracleCloth(d, dict(x=sphere_tm.x, faces=sphere_tm.faces), dhat=sim.dHat,
                           stiffness_stretch=sim.stiffness_stretch, stiffness_bending=sim.stiffness_bending,
                           g=sim.g, dt=sim.dt, padding=sim.padding, collision=True, bvh="tri",
                           cap=sim.cache_size, refit_every_substep=True)


def test_batched_tree_subtrees_match_single_trees(slx, orc, rng):
    import torch
    import importlib
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    batch, cloths, spheres, bs = small_batch(slx)
    tree = bp.DeviceTree(bs.cloth, "tri")
    pad = 0.02
    tree.build(pad)
    q = bs.cloth.x4.clone()
    q[:, :3] += torch.randn_like(q[:, :3]) * 0.01
    cand, cnt = tree.query(q, pad, 40, q_inst=bs.cloth.inst["vert_inst"])
    cand, cnt = cand.cpu().numpy(), cnt.cpu().numpy()
    vb, fb = bs.cloth.inst_np["vert_base"], bs.cloth.inst_np["face_base"]
    qh = q[:, :3].cpu().numpy()
    for i, c in enumerate(cloths):
        ot = orc.TriBVH(c.faces, c.x, pad)
        ql = qh[vb[i]:vb[i + 1]]
        oc, on, _ = ot.query(ql - np.float32(pad), ql + np.float32(pad), 40)
        np.testing.assert_array_equal(cnt[vb[i]:vb[i + 1]], on)
        for k in range(len(ql)):
            got = cand[:on[k], vb[i] + k] - fb[i]
            np.testing.assert_array_equal(got, oc[k, :on[k]])


def test_batched_frames_equal_independent_instances(slx, orc):
    batch, cloths, spheres, bs = small_batch(slx)
    vb = bs.cloth.inst_np["vert_base"]
    fixed = bs.cloth.fixed.cpu().numpy()
    refs = [oracle_instance(orc, c, s, bs, fixed[vb[i]:vb[i + 1]].copy()) for i, (c, s) in
            enumerate(zip(cloths, spheres))]
    for _ in range(6):
        bs.forward(10, 2)
        for r in refs:
            r.forward(10, 2)
    x = bs.cloth.positions()
    for i, r in enumerate(refs):
        np.testing.assert_array_equal(x[vb[i]:vb[i + 1]], r.x)
    assert bs.sim.status() == 0


def test_graph_replay_equals_eager(slx):
    import torch
    batch, cloths, spheres, bs = small_batch(slx, seed=5)
    _, _, _, eager = small_batch(slx, seed=5)
    bs.capture(10, 2)
    eager.forward(10, 2)            # the capture ran one eager warm-up frame
    for k in range(3):
        bs.replay()
        eager.forward(10, 2)
        torch.cuda.synchronize()
        _ = (bs.cloth.x4 * 2).sum().item()      # unrelated eager work between replays
        np.testing.assert_array_equal(bs.cloth.positions(), eager.cloth.positions())


def test_launch_shape_does_not_change_results(slx):
    """Per-solver launch shapes (spring kernels, walk stack, active grid, refit
    shape), and the library default set process-wide, give the same frames."""
    A = __import__("importlib").import_module("starlab-physics_amd._abi")
    outs = []
    for shape, walk, grid, refit in (((256, 1), None, None, None), ((64, 4), 16, 3, (1024, 512)),
                                     ((128, 2), 64, 1, (2048, 2048))):
        _, _, _, bs = small_batch(slx, seed=7)
        bs.sim.spring_launch, bs.sim.walk_stack, bs.sim.active_grid, bs.sim.refit_shape = shape, walk, grid, refit
        for _ in range(3):
            bs.forward(10, 2)
        outs.append(bs.cloth.positions())
    A.call("slx_set_spring_launch", 64, 2)
    _, _, _, bs = small_batch(slx, seed=7)
    for _ in range(3):
        bs.forward(10, 2)
    outs.append(bs.cloth.positions())
    A.call("slx_set_spring_launch", 256, 1)
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)


def test_two_solvers_keep_their_own_launch_shapes(slx):
    """Two scenes with different launch shapes interleaved frame by frame equal
    the same scenes run one after the other."""
    runs = []
    for interleave in (True, False):
        _, _, _, a = small_batch(slx, seed=11)
        _, _, _, b = small_batch(slx, seed=11)
        a.sim.spring_launch, b.sim.spring_launch = (64, 4), (256, 2)
        a.sim.walk_stack, b.sim.walk_stack = 16, 64
        order = [a, b] * 3 if interleave else [a] * 3 + [b] * 3
        for s in order:
            s.forward(10, 2)
        runs.append((a.cloth.positions(), b.cloth.positions()))
        assert a.sim._ctx != b.sim._ctx
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    np.testing.assert_array_equal(runs[0][0], runs[0][1])


def test_autotune_picks_a_valid_shape(slx):
    """The tuner keeps the fastest spring shape and a best value of each other
    knob; the tuned scene's frames equal an untuned scene's."""
    batch, cloths, spheres, bs = small_batch(slx, seed=9)
    best, res = bs.autotune(5, 1, reps=2)
    assert best in res and len(res) == 9
    assert bs.sim.spring_launch == best
    assert bs.tuning["walk_stack"]["best"] in (16, 24, 32, 64) and bs.sim.walk_stack == bs.tuning["walk_stack"]["best"]
    assert bs.tuning["cache_margin"]["best"] in (0.125, 0.25, 0.5)
    assert bs.tuning["refit_sub"]["best"] in (1024, 2048)
    assert bs.tuning["static_cache_every"]["best"] in (1, 4)
    assert bs.sim.static_cache_every == bs.tuning["static_cache_every"]["best"]
    _, _, _, ref = small_batch(slx, seed=9)
    ref.forward(5, 1)                 # the final capture ran one eager frame from the initial state
    for _ in range(3):
        bs.replay()
        ref.forward(5, 1)
    np.testing.assert_array_equal(bs.cloth.positions(), ref.cloth.positions())
