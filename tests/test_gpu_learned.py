"""GPU checks of the learned warm start and the residual solve it seeds: the
residual-solve passes bit for bit against the oracle (iterates and residuals)
with their self-terminating chain; features and MLP forward against the float64
reference (BF16 operands emulated; tolerance 2e-2 relative on outputs, 1e-5 on
features); gradients of the MFMA backward against autograd (relative error of
the whole gradient <= 3e-2); training reduces the loss; the trained network
reaches the residual in fewer Jacobi iterations than plain Jacobi on a held-out
seed; a broken network falls back to the plain solve with identical results."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def L():
    return importlib.import_module("starlab-physics_amd.learned")


def cloth(slx, n=48, seed=0):
    """Cloth pinned at two corners with stiff stretch: the reference's k_s slot
    (stretch, on y) gets stiffness_bending * dt^2 (XPBD.py:354-355), so both
    stiffnesses are 1e6 here."""
    import torch
    g = slx.scene.grid_cloth(n, size=1.0, height=0.5, jitter=1e-3, seed=seed)
    mesh = slx.Mesh(x=g.x, faces=g.faces)
    sim = slx.Solver(mesh, None, dHat=1e-4, stiffness_stretch=1e6, stiffness_bending=1e6,
                     g=(0.0, -9.8, 0.0), dt=1.0 / 60.0)
    sel = torch.zeros(mesh.num_verts, device="cuda")
    sel[0] = 1
    sel[n - 1] = 1
    sim.set_fixed_vertices(sel)
    return mesh, sim


def test_implicit_passes_bit_exact_and_self_terminating(slx):
    import torch
    from oracle import oracle as orc
    learned = L()
    mesh, sim = cloth(slx, 24)
    rs = learned.ResidualSolver(sim)
    for _ in range(12):                       # a few plain substeps: the cloth swings
        rs.substep(10, tol=1e-5, max_iters=400)
    s = rs._scalars(10)
    rs.reference_pass(s)
    y1 = rs.y1[:, :3].cpu().numpy().copy()
    fixed = mesh.fixed.cpu().numpy()
    hd = mesh.host_dict()
    npass = 6
    oy = y1.copy()
    ores = [orc.implicit_pass(hd, y1, oy, fixed, hd["m_inv"], np.float32(rs.kappa)) for _ in range(npass)]
    A = importlib.import_module("starlab-physics_amd._abi")
    res = torch.zeros(npass + 1, dtype=torch.int32, device="cuda")
    conv = torch.zeros(1, dtype=torch.int32, device="cuda")
    rs.ya.copy_(rs.y1)
    A.call("slx_solve_reset", npass + 1, A.ptr(res), A.ptr(conv), A.stream_handle())
    A.call("slx_implicit_passes", mesh.num_verts, A.ptr(mesh.fixed), A.ptr(rs.y1), A.ptr(rs.ya), A.ptr(rs.yb),
           A.ptr(mesh.adj_slice_off), A.ptr(mesh.adj_slice_width), A.ptr(mesh.adj_entries), rs.kappa, 0, npass, 0.0,
           A.ptr(res), A.ptr(conv), A.stream_handle())
    gres = res.cpu().numpy().view(np.float32)[:npass]
    np.testing.assert_array_equal(gres, np.array(ores, np.float32))
    out = rs.ya if npass % 2 == 0 else rs.yb
    np.testing.assert_array_equal(out[:, :3].cpu().numpy(), oy)
    assert ores[0] > 1e-7 and ores[-1] < ores[0]
    # a chain launched past convergence stops by itself: tol between r_1 and r_2
    tol = float(np.float32(ores[2]))
    first = next(k for k, r in enumerate(ores) if r <= tol)
    rs.ya.copy_(rs.y1)
    A.call("slx_solve_reset", npass + 1, A.ptr(res), A.ptr(conv), A.stream_handle())
    A.call("slx_implicit_passes", mesh.num_verts, A.ptr(mesh.fixed), A.ptr(rs.y1), A.ptr(rs.ya), A.ptr(rs.yb),
           A.ptr(mesh.adj_slice_off), A.ptr(mesh.adj_slice_width), A.ptr(mesh.adj_entries), rs.kappa, 0, npass, tol,
           A.ptr(res), A.ptr(conv), A.stream_handle())
    gres = res.cpu().numpy().view(np.float32)
    assert conv.item() == first
    np.testing.assert_array_equal(gres[:first + 1], np.array(ores[:first + 1], np.float32))
    assert np.all(gres[first + 1:] == 0)
    # y_first is intact in its buffer
    oy = y1.copy()
    for _ in range(first):
        orc.implicit_pass(hd, y1, oy, fixed, hd["m_inv"], np.float32(rs.kappa))
    got = rs.yb if first & 1 else rs.ya
    np.testing.assert_array_equal(got[:, :3].cpu().numpy(), oy)


def test_mfma_probe_layout():
    """The MLP's MFMA operand layout assumption, on a known product."""
    import torch
    learned = L()
    md = learned.WarmStart(seed=1)
    rng = np.random.default_rng(0)
    feat = torch.from_numpy(rng.normal(0, 1, (96, 24)).astype(np.float32)).cuda()
    tgt = torch.zeros((96, 3), dtype=torch.float32, device="cuda")
    g, loss = md.grad(feat, tgt)
    from oracle import mlp_ref
    out, _ = mlp_ref.forward(mlp_ref.unpack(md.w.cpu().numpy()), feat.cpu().numpy(), np.zeros(24), np.ones(24))
    ref_loss = float((out[:, :3] ** 2).sum() / 96)
    assert abs(loss.item() - ref_loss) <= 2e-2 * ref_loss


def test_features_match_reference(slx):
    import torch
    from oracle import mlp_ref
    mesh, sim = cloth(slx, 24)
    for _ in range(3):
        sim.forward(10, 1)
    rs = L().ResidualSolver(sim)
    s = rs._scalars(10)
    rs.reference_pass(s)
    F = rs.features(s).cpu().numpy()
    ref = mlp_ref.features(mesh.positions(), mesh.v4[:, :3].cpu().numpy(), rs.y1[:, :3].cpu().numpy(),
                           mesh.adj, s["dt"], rs.inv_s)
    np.testing.assert_allclose(F, ref, rtol=1e-4, atol=1e-5)
    assert np.abs(ref).max() > 1e-3


def test_gradient_matches_autograd():
    import torch
    from oracle import mlp_ref
    learned = L()
    md = learned.WarmStart(seed=2)
    rng = np.random.default_rng(1)
    n = 1000
    feat = rng.normal(0, 1, (n, 24)).astype(np.float32)
    tgt = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    fm, fi = rng.normal(0, 0.1, 24).astype(np.float32), rng.uniform(0.8, 1.2, 24).astype(np.float32)
    md.fmean.copy_(torch.from_numpy(fm))
    md.fistd.copy_(torch.from_numpy(fi))
    g, loss = md.grad(torch.from_numpy(feat).cuda(), torch.from_numpy(tgt).cuda())
    L_ref, g_ref = mlp_ref.loss_and_grad(md.w.cpu().numpy(), feat, tgt, fm, fi)
    g = g.cpu().numpy()
    assert abs(loss.item() - L_ref) <= 1e-2 * L_ref
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"MFMA backward vs float64 autograd: relative error {rel:.3e}")
    assert rel < 3e-2
    # per layer, and padded parameters have zero gradient
    o = 0
    for name, shp in mlp_ref.LAYOUT:
        k = int(np.prod(shp))
        a, b = g[o:o + k], g_ref[o:o + k]
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7, name
        o += k
    P = mlp_ref.unpack(g)
    assert np.all(P["W1"][:, 24:].numpy() == 0) and np.all(P["W4"][3:].numpy() == 0)


def test_forward_matches_reference_through_warm_start(slx):
    import torch
    from oracle import mlp_ref
    mesh, sim = cloth(slx, 20)
    for _ in range(3):
        sim.forward(10, 1)
    learned = L()
    md = learned.WarmStart(seed=3)
    rs = learned.ResidualSolver(sim, md)
    s = rs._scalars(10)
    rs.reference_pass(s)
    F = rs.features(s).cpu().numpy()
    # normalised inputs of order one, and an output scale whose corrections are
    # far above float32 resolution of the positions
    md.fmean.copy_(torch.from_numpy(F.mean(0).astype(np.float32)))
    sd = F.std(0)
    md.fistd.copy_(torch.from_numpy(np.where(sd > 0, 1.0 / np.maximum(sd, 1e-30), 0.0).astype(np.float32)))
    md.out_scale = 1e-2
    rs.warm(s)
    got = (rs.ya[:, :3] - rs.y1[:, :3]).cpu().numpy() / md.out_scale
    out, _ = mlp_ref.forward(mlp_ref.unpack(md.w.cpu().numpy()), F, md.fmean.cpu().numpy(), md.fistd.cpu().numpy())
    ref = out[:, :3].detach().numpy() * mesh.fixed.cpu().numpy()[:, None]
    assert np.abs(ref).max() > 1e-2
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-2, err


def test_training_and_warm_start_gain(slx):
    """Train on plain solves (residual 1e-6) of seeded variants; on a held-out
    seed the warm start reaches the residual 1e-5 in fewer Jacobi iterations, at a
    wall-time cost bounded by its inference launch (32 x 32 cloth: the passes it
    saves cost less than that launch; the benchmark's time-to-residual line records
    the 256 x 256 cloth)."""
    import torch
    learned = L()
    rng = np.random.default_rng(0)
    feats, tgts = [], []
    for seed in (1, 2, 3):
        mesh, sim = cloth(slx, 32, seed)
        rs = learned.ResidualSolver(sim)
        for _ in range(40):
            rs.collect(10, 1e-6, 2000, 1000, rng, feats, tgts)
    feat = torch.cat(feats).contiguous()
    tgt = torch.cat(tgts).contiguous()
    assert torch.isfinite(feat).all() and torch.isfinite(tgt).all()
    md = learned.WarmStart(seed=0)
    md.set_normalisation(feat, tgt)
    zero = float(((tgt / md.out_scale) ** 2).sum(dim=1).mean().item())    # the loss of predicting 0
    losses = md.train(feat, tgt, epochs=30, batch=4096, lr=2e-3)
    print("training loss first / last epoch, zero-prediction loss:", losses[0], losses[-1], zero)
    assert np.isfinite(losses).all() and losses[-1] < 0.3 * losses[0] and losses[-1] < 0.5 * zero
    import time
    res = {}
    for learned_on in (False, True):
        mesh, sim = cloth(slx, 32, seed=7)
        rs = learned.ResidualSolver(sim, md)
        rs.substep(10, tol=1e-5, max_iters=2000, learned=learned_on)     # first launches of each kernel
        torch.cuda.synchronize()
        it, falls = 0, 0
        t0 = time.perf_counter()
        for _ in range(30):
            k, fell = rs.substep(10, tol=1e-5, max_iters=2000, learned=learned_on)
            it += k
            falls += fell
        torch.cuda.synchronize()
        res[learned_on] = (it, falls, time.perf_counter() - t0)
    print("iterations, fallbacks, wall s: plain", res[False], "learned", res[True])
    assert res[True][0] < 0.5 * res[False][0]
    # wall time: on this small cloth a pass costs less than the network's one
    # launch, so the warm start saves passes but not time; the bound keeps its
    # cost to that launch (measured: plain 2.7 ms, learned 3.3 ms per 30 substeps)
    assert res[True][2] < 1.5 * res[False][2]


def test_fallback_when_the_network_is_wrong(slx):
    learned = L()
    mesh, sim = cloth(slx, 24)
    md = learned.WarmStart(seed=5)
    md.out_scale = float("nan")   # a broken network: the residual is not finite
    rs = learned.ResidualSolver(sim, md)
    k, fell = rs.substep(10, tol=1e-5, max_iters=2000, learned=True)
    assert fell
    assert np.isfinite(mesh.positions()).all()
    # the fallback substep equals a plain substep from the same state
    mesh2, sim2 = cloth(slx, 24)
    rs2 = learned.ResidualSolver(sim2)
    rs2.substep(10, tol=1e-5, max_iters=2000, learned=False)
    np.testing.assert_array_equal(mesh.positions(), mesh2.positions())
