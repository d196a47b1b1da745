"""GPU parity of the edge-edge row against the oracle, bit for bit on the same
float32 inputs: edge-box candidate lists of both tree kinds (per-lane and
wavefront walks), the position pass with the edge-edge constraints after the
vertex-triangle ones (deltas, counts, pair records), and the velocity pass over
those records.

The scene is coarse on purpose: the reference's region rule only reaches the
interior type when (u x v).(x0 - x2) >= 1e-4 (u, v the edge directions), a
dimensional bound that fine meshes (edges of a few hundredths) never meet.  Two
dynamic grids with edges of about 0.27, rotated against each other and a small
gap apart, and a static grid below, cross in many places within sqrt(dhat)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DHAT = 1e-2


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def abi():
    import importlib
    return importlib.import_module("starlab-physics_amd._abi")


def rot_y(x, deg):
    t = np.deg2rad(deg)
    R = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    return x @ R.T


def ee_scene(slx, n=16, size=4.0, gap=0.05, seed=0):
    """Dynamic mesh: two n x n grids (the upper rotated 30 degrees about y, gap
    above); static mesh: a grid rotated 60 degrees, gap below the lower one."""
    rng = np.random.default_rng(seed)
    a = slx.scene.grid_cloth(n, size=size, height=0.0, jitter=0.01, seed=seed)
    b = slx.scene.grid_cloth(n, size=size, height=0.0, jitter=0.01, seed=seed + 1)
    xa = a.x.astype(np.float64)
    xb = rot_y(b.x.astype(np.float64), 30.0)
    xb[:, 1] += gap
    xa[:, 1] += rng.uniform(-0.01, 0.01, len(xa))
    x = np.concatenate([xa, xb]).astype(np.float32)
    faces = np.concatenate([a.faces, b.faces + len(xa)]).astype(np.int32)
    cloth = slx.Mesh(x=x, faces=faces)
    c = slx.scene.grid_cloth(n, size=size, height=0.0, jitter=0.01, seed=seed + 2)
    xc = rot_y(c.x.astype(np.float64), 60.0)
    xc[:, 1] -= gap
    obst = slx.Mesh(x=xc.astype(np.float32), faces=c.faces, is_static=True)
    return cloth, obst


def edge_boxes(x, edges, pad):
    p0, p1 = x[edges[:, 0]], x[edges[:, 1]]
    return np.minimum(p0, p1) - np.float32(pad), np.maximum(p0, p1) + np.float32(pad)


def traverse_edges(A, tree, edges_dev, x4, ne, pad, cap, order=None, redo=None):
    import torch
    cand = torch.full((cap * max(ne, 1),), -1, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(max(ne, 1), dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = A.ptr
    cell = tree.mode == "cell"
    A.call("slx_traverse_edges", 0 if cell else 1, ne, P(edges_dev), P(x4), float(np.float32(pad)),
           tree.n_leaves, P(tree.num) if cell else None, P(tree.sorted_faces), P(tree.nlo4), P(tree.nhi4),
           None if cell else P(tree.quad4), P(tree.fmin4) if cell else None, P(tree.fmax4) if cell else None,
           P(tree.box6) if cell else None, cap, P(cand), P(cnt), P(st), P(order), P(redo), A.stream_handle())
    return cand.view(cap, -1)[:, :ne], cnt[:ne], st


@pytest.mark.parametrize("mode,pad,cap", [("cell", 0.05, 40), ("tri", 0.05, 40), ("tri", 0.2, 6)])
def test_edge_traversal_bit_exact(slx, orc, mode, pad, cap):
    import importlib
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    A = abi()
    cloth, obst = ee_scene(slx)
    for mesh_q, mesh_t in ((cloth, obst), (cloth, cloth)):
        tree = bp.DeviceTree(mesh_t, mode)
        tree.build(pad)
        Ot = orc.CellBVH if mode == "cell" else orc.TriBVH
        ot = Ot(mesh_t.fid_np, mesh_t.positions(), pad)
        ne = mesh_q.num_edges
        print(f"[ee traverse] mode={mode} pad={pad} cap={cap} ne={ne} leaves={tree.n_leaves}", flush=True)
        cand, cnt, st = traverse_edges(A, tree, mesh_q.edge_indices, mesh_q.x4, ne, pad, cap)
        qmin, qmax = edge_boxes(mesh_q.positions(), mesh_q.eid_np, pad)
        oc, on, fl = ot.query(qmin, qmax, cap)
        cnt = cnt.cpu().numpy()
        cand = cand.cpu().numpy()
        np.testing.assert_array_equal(cnt, on)
        for i in range(ne):
            np.testing.assert_array_equal(cand[:cnt[i], i], oc[i, :on[i]])
        assert (st.item() != 0) == (fl > 0)
        assert on.sum() > ne


@pytest.mark.parametrize("depth", [16, 32])
def test_edge_traversal_two_tier_bit_exact(slx, orc, depth):
    """The two-tier walk over edge boxes, in the Morton-rank edge order: a
    shallow first-tier stack (16 entries overflow here, for thousands of edges at
    padding 0.5), the overflowing edges walked again in full; the oracle's lists."""
    import importlib
    import torch
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    A = abi()
    cloth, _ = ee_scene(slx, n=40)
    pad, cap = 0.5, 64
    tree = bp.DeviceTree(cloth, "tri")
    tree.build(pad)
    ne = cloth.num_edges
    redo = torch.zeros(1 + ne, dtype=torch.int32, device="cuda")
    print(f"[ee two-tier] depth={depth} ne={ne}", flush=True)
    A.call("slx_set_walk_stack", depth)
    try:
        cand, cnt, st = traverse_edges(A, tree, cloth.edge_indices, cloth.x4, ne, pad, cap,
                                       order=cloth.edge_order(), redo=redo)
    finally:
        A.call("slx_set_walk_stack", 24)   # the library default
    ot = orc.TriBVH(cloth.fid_np, cloth.positions(), pad)
    oc, on, fl = ot.query(*edge_boxes(cloth.positions(), cloth.eid_np, pad), cap)
    cnt, cand = cnt.cpu().numpy(), cand.cpu().numpy()
    np.testing.assert_array_equal(cnt, on)
    for i in range(ne):
        np.testing.assert_array_equal(cand[:cnt[i], i], oc[i, :on[i]])
    assert (st.item() != 0) == (fl > 0)
    if depth == 16:
        assert redo[0].item() > 0, "the shallow stack must overflow for some edge query"


def test_edge_traversal_wavefront_walk(slx, orc):
    """Many edges against a coarse static tree (leaves x 8 <= queries): the
    wavefront walk serves the edge boxes with the same lists."""
    import importlib
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    A = abi()
    cloth, _ = ee_scene(slx, n=40)
    c = slx.scene.grid_cloth(6, size=4.0, height=-0.05, jitter=0.01, seed=9)
    obst = slx.Mesh(x=rot_y(c.x.astype(np.float64), 45.0).astype(np.float32), faces=c.faces, is_static=True)
    pad, cap = 0.05, 40
    tree = bp.DeviceTree(obst, "tri")
    tree.build(pad)
    assert tree.n_leaves * 8 <= cloth.num_edges
    cand, cnt, st = traverse_edges(A, tree, cloth.edge_indices, cloth.x4, cloth.num_edges, pad, cap)
    ot = orc.TriBVH(obst.fid_np, obst.positions(), pad)
    oc, on, fl = ot.query(*edge_boxes(cloth.positions(), cloth.eid_np, pad), cap)
    cnt, cand = cnt.cpu().numpy(), cand.cpu().numpy()
    np.testing.assert_array_equal(cnt, on)
    for i in range(cloth.num_edges):
        np.testing.assert_array_equal(cand[:cnt[i], i], oc[i, :on[i]])
    assert on.sum() > cloth.num_edges


def ee_pass(slx, orc, *, pcap=None, rec_cap=128, seed=0, velocity=True):
    """One position pass (and velocity pass) on the GPU and in the oracle, on the
    same y and the same candidate lists."""
    import importlib
    import torch
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    A = abi()
    P = A.ptr
    rng = np.random.default_rng(seed)
    cloth, obst = ee_scene(slx, seed=seed)
    pad, cap = 0.06, 40
    nv, nvs, ne = cloth.num_verts, obst.num_verts, cloth.num_edges
    pcap = 3 * cap if pcap is None else pcap
    y = cloth.positions() + rng.normal(0, 2e-3, (nv, 3)).astype(np.float32)
    cloth.y4[:, :3] = T(y)
    fixed = np.ones(nv, np.float32)
    fixed[:2] = 0
    cloth.fixed.copy_(T(fixed))
    cloth._refresh_products()
    td, ts = bp.DeviceTree(cloth, "tri"), bp.DeviceTree(obst, "tri")
    td.build(pad)
    ts.build(pad)
    c0, n0 = ts.query(cloth.y4, pad, cap)
    c1, n1 = td.query(cloth.y4, pad, cap)
    c2, n2 = td.query(obst.x4, pad, cap)
    ce_st, ne_st, _ = traverse_edges(A, ts, cloth.edge_indices, cloth.y4, ne, pad, cap)
    ce_dy, ne_dy, _ = traverse_edges(A, td, cloth.edge_indices, cloth.y4, ne, pad, cap)
    ce_st, ce_dy = ce_st.contiguous(), ce_dy.contiguous()
    nqt = 2 * nv + nvs
    ws = torch.empty(A.lib().slx_collision_ws_bytes(nv, nvs, rec_cap), dtype=torch.uint8, device="cuda")
    pn = torch.zeros(nqt, dtype=torch.int32, device="cuda")
    pcnt = torch.zeros(nqt, dtype=torch.int32, device="cuda")
    prec = torch.zeros((cap * 4 * nqt, 4), dtype=torch.float32, device="cuda")
    epn = torch.zeros(2 * ne, dtype=torch.int32, device="cuda")
    epc = torch.zeros(2 * ne, dtype=torch.int32, device="cuda")
    eprec = torch.zeros((max(pcap, 1) * 4 * 2 * ne, 4), dtype=torch.float32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    fe_dy, fe_st = cloth.face_edge_indices(), obst.face_edge_indices()
    E = A.EdgeSets(ne, P(cloth.edge_indices), P(obst.edge_indices), P(fe_dy), P(fe_st), cap, P(ce_st), P(ne_st),
                   P(ce_dy), P(ne_dy), pcap, P(epn), P(epc), P(eprec))
    print(f"[ee pass] nv={nv} ne={ne} pcap={pcap} rec_cap={rec_cap}", flush=True)
    A.call("slx_collision_pass_ee", nv, nvs, P(cloth.face_indices), P(obst.face_indices), P(cloth.y4), P(obst.x4),
           P(cloth.fixed), P(cloth.m_inv), float(np.float32(DHAT)), float(np.float32(1e8)), cap, P(c0), P(n0),
           P(c1), P(n1), P(c2), P(n2), P(pn), cap, P(pcnt), P(prec), A.C.byref(E), rec_cap, P(st), P(ws),
           ws.numel(), A.stream_handle())
    to_host = lambda c, n: (np.ascontiguousarray(c.cpu().numpy().T), n.cpu().numpy())
    yo = y.copy()
    ee = dict(edges_dy=cloth.eid_np, edges_st=obst.eid_np, fe_dy=cloth.fe_np, fe_st=obst.fe_np,
              st=to_host(ce_st, ne_st), dy=to_host(ce_dy, ne_dy), pcap=pcap)
    out = orc.collision_pass(dict(faces=cloth.fid_np), obst.fid_np, yo, obst.positions(), fixed, cloth.m_inv_np,
                             np.float32(DHAT), np.float32(1e8), cap, to_host(c0, n0), to_host(c1, n1),
                             to_host(c2, n2), record=True, ee=ee)
    res = dict(cloth=cloth, obst=obst, y=y, yo=yo, out=out, st=st, epn=epn, epc=epc, eprec=eprec, ne=ne,
               pcap=pcap, fixed=fixed)
    if not velocity:
        return res
    v = rng.normal(0, 0.5, (nv, 3)).astype(np.float32)
    cloth.v4[:, :3] = T(v)
    vs = np.zeros((nvs, 3), np.float32)
    vs[:, 1] = rng.uniform(-0.5, 0.5, nvs).astype(np.float32)
    vs4 = torch.zeros((nvs, 4), dtype=torch.float32, device="cuda")
    vs4[:, :3] = T(vs)
    A.call("slx_velocity_pass_ee", nv, nvs, P(cloth.face_indices), P(obst.face_indices), P(cloth.v4), P(vs4),
           P(cloth.fixed), P(cloth.m_inv), P(obst.m_inv), float(np.float32(0.8)), cap, P(pcnt), P(prec),
           A.C.byref(E), None, 0.0, rec_cap, P(st), P(ws), ws.numel(), A.stream_handle())
    vo = v.copy()
    orc.velocity_pass(dict(faces=cloth.fid_np), obst.fid_np, vo, vs, fixed, cloth.m_inv_np, obst.m_inv_np,
                      np.float32(0.8), out["pairs"], cap, ee=out["ee"])
    res.update(v=v, vo=vo)
    return res


def test_collision_pass_ee_bit_exact(slx, orc):
    r = ee_pass(slx, orc, velocity=False)
    E, ne = r["out"]["ee"], r["ne"]
    assert r["st"].item() == 0
    assert E["pn_st"].sum() > 20 and E["pn_dy"].sum() > 20, "scene must produce edge-edge contacts"
    np.testing.assert_array_equal(r["cloth"].y4[:, :3].cpu().numpy(), r["yo"])
    epn = r["epn"].cpu().numpy()
    np.testing.assert_array_equal(epn[:ne], E["pn_st"])
    np.testing.assert_array_equal(epn[ne:], E["pn_dy"])
    # pair records: column-major float4 [(j*4 + w) * 2ne + q]
    epc = r["epc"].cpu().numpy()
    rec = r["eprec"].cpu().numpy().reshape(r["pcap"], 4, 2 * ne, 4)
    for s, nm in enumerate(("st", "dy")):
        np.testing.assert_array_equal(epc[s * ne:(s + 1) * ne], np.minimum(E["pn_" + nm], r["pcap"]))
        for e in np.nonzero(E["pn_" + nm])[0]:
            for j in range(min(E["pn_" + nm][e], r["pcap"])):
                q = s * ne + e
                g = rec[j, :, q, :3]
                np.testing.assert_array_equal(g, E["pg_" + nm][e, j].astype(np.float32))
                assert rec[j, 0, q, 3] == np.float32(E["ps_" + nm][e, j])
                assert rec[j, 1, q, 3:].view(np.int32)[0] == E["pi_" + nm][e, j, 0]
                assert rec[j, 2, q, 3:].view(np.int32)[0] == E["pi_" + nm][e, j, 1] == 8


def test_collision_pass_ee_pair_overflow(slx, orc):
    """pair_cap 1: the counts stay whole, one record is kept per edge and set,
    the status says so; the positions are unaffected."""
    r = ee_pass(slx, orc, pcap=1, velocity=False)
    E, ne = r["out"]["ee"], r["ne"]
    np.testing.assert_array_equal(r["cloth"].y4[:, :3].cpu().numpy(), r["yo"])
    np.testing.assert_array_equal(r["epc"].cpu().numpy(), np.minimum(np.concatenate([E["pn_st"], E["pn_dy"]]), 1))
    assert (np.concatenate([E["pn_st"], E["pn_dy"]]) > 1).any()
    assert r["st"].item() & 16


def test_velocity_pass_ee_bit_exact(slx, orc):
    r = ee_pass(slx, orc)
    assert r["st"].item() == 0
    got = r["cloth"].v4[:, :3].cpu().numpy()
    assert np.any(got != r["v"])
    np.testing.assert_array_equal(got, r["vo"])


def ee_frame_scene(slx, mode, velocity, padding=0.06, n=12):
    """ee_scene's meshes 0.11 apart (just outside sqrt(dhat)) under gravity, two
    corners of the lower grid pinned, the edge-edge constraints on: the lower
    grid sags onto the static one and the upper grid onto it; the oracle driver
    on the same inputs."""
    import torch
    from oracle import oracle as orc
    cloth, obst = ee_scene(slx, n=n, gap=0.11)
    sim = slx.Solver(cloth, obst, dHat=DHAT, stiffness_stretch=1e4, stiffness_bending=1e2,
                     g=(0.0, -9.8, 0.0), dt=1.0 / 60.0)
    sim.enable_collision_handling = True
    sim.enable_edge_edge = True
    sim.enable_velocity_update = velocity
    sim.bvh_mode = mode
    sim.padding = padding
    sim.records_per_vertex = 128
    sel = torch.zeros(cloth.num_verts, device="cuda")
    sel[0] = 1
    sel[n - 1] = 1
    sim.set_fixed_vertices(sel)
    st = dict(x=obst.positions(), faces=obst.fid_np, m_inv=obst.m_inv_np, edges=obst.eid_np,
              fe=slx.topology.face_edges(obst.fid_np, obst.eid_np))
    ref = orc.OracleCloth(cloth.host_dict(), st, dhat=sim.dHat, stiffness_stretch=sim.stiffness_stretch,
                          stiffness_bending=sim.stiffness_bending, g=sim.g, dt=sim.dt, padding=sim.padding,
                          collision=True, bvh=mode, cap=sim.cache_size, velocity=velocity, mu=sim.mu,
                          edge_edge=True)
    return cloth, sim, ref


@pytest.mark.parametrize("mode,velocity", [("tri", False), ("tri", True), ("cell", True)])
def test_frames_with_edge_edge_bit_exact(slx, mode, velocity):
    cloth, sim, ref = ee_frame_scene(slx, mode, velocity)
    pairs = 0
    for frame in range(6):
        print(f"[ee frame] mode={mode} velocity={velocity} frame={frame}", flush=True)
        before = ref.stats["overflow"]
        sim.clear_status()
        sim.forward(4)
        ref.forward(4)
        st = sim.status()
        assert st & ~1 == 0, (frame, st)
        assert (st & 1 != 0) == (ref.stats["overflow"] > before), (frame, st)
        np.testing.assert_array_equal(cloth.positions(), ref.x)
        e = sim.edge_pairs()
        num = e["num"].cpu().numpy()
        ne = cloth.num_edges
        np.testing.assert_array_equal(num[:ne], ref.last_ee["pn_st"])
        np.testing.assert_array_equal(num[ne:], ref.last_ee["pn_dy"])
        pairs += int(num.sum())
    assert pairs > 50, "the layers must touch"
    if velocity:
        np.testing.assert_array_equal(cloth.v4[:, :3].cpu().numpy(), ref.v.astype(np.float32))


def test_frames_with_edge_edge_shallow_walk(slx):
    """A 16-entry first-tier stack, a deeper tree (two 40 x 40 grids) and large
    boxes (padding 0.3, where the standalone walk above overflows it): edge
    queries the first tier cannot serve go through the redo list and the full
    walk; the frames stay bit-exact."""
    A = abi()
    cloth, sim, ref = ee_frame_scene(slx, "tri", True, padding=0.3, n=40)
    A.call("slx_set_walk_stack", 16)
    try:
        for frame in range(3):
            print(f"[ee frame shallow] frame={frame}", flush=True)
            sim.forward(4)
            ref.forward(4)
            np.testing.assert_array_equal(cloth.positions(), ref.x)
    finally:
        A.call("slx_set_walk_stack", 24)   # the library default


@pytest.mark.parametrize("blocks", [1, 3])
def test_device_count_walks_loop(slx, orc, blocks):
    """Walks whose query count is on the device (the second tier's redo lists,
    the edge gate's live lists) run on a capped grid and loop over their
    queries; with one or three workgroups every such walk takes many
    iterations, and the lists and the frames stay bit-exact."""
    A = abi()
    print(f"[active grid] blocks={blocks}", flush=True)
    A.call("slx_set_active_grid", blocks)
    try:
        test_edge_traversal_two_tier_bit_exact(slx, orc, 16)
        test_frames_with_edge_edge_shallow_walk(slx)
    finally:
        A.call("slx_set_active_grid", 2048)


@pytest.mark.parametrize("velocity", [False, True])
def test_frames_with_edge_gate_on_fine_cloth(slx, velocity):
    """A 64 x 64 cloth (edges of 1.6-2.2 cm) falling on a sphere, padding 0.02: the
    edge gate's bound |u| LY (|u| + sqrt(3) (LX + 2 pad + D)) stays below 1e-4
    for every cloth edge while the cloth falls freely, so the dynamic edge set is
    skipped whole; once the sphere pushes vertices, D (the pass's largest y - x)
    lets part of the set through.  Every frame equals the oracle's, which runs
    every edge-edge test."""
    import torch
    from oracle import oracle as orc
    g = slx.scene.grid_cloth(64, size=1.0, height=0.35, jitter=1e-3, seed=0)
    cloth = slx.Mesh(x=g.x, faces=g.faces)
    ico = slx.scene.icosphere(3, 0.3)
    obst = slx.Mesh(x=ico.x, faces=ico.faces, is_static=True)
    sim = slx.Solver(cloth, obst, dHat=1e-4, stiffness_stretch=1e6, stiffness_bending=1e3,
                     g=(0.0, -9.8, 0.0), dt=1.0 / 60.0)
    sim.enable_collision_handling = True
    sim.enable_edge_edge = True
    sim.enable_velocity_update = velocity
    sim.bvh_mode = "tri"
    sim.refit_every_substep = True
    sim.padding = 0.02
    sel = torch.zeros(cloth.num_verts, device="cuda")
    sel[0] = 1
    sel[63] = 1
    sim.set_fixed_vertices(sel)
    st = dict(x=obst.positions(), faces=obst.fid_np, m_inv=obst.m_inv_np, edges=obst.eid_np,
              fe=slx.topology.face_edges(obst.fid_np, obst.eid_np))
    ref = orc.OracleCloth(cloth.host_dict(), st, dhat=sim.dHat, stiffness_stretch=sim.stiffness_stretch,
                          stiffness_bending=sim.stiffness_bending, g=sim.g, dt=sim.dt, padding=sim.padding,
                          collision=True, bvh="tri", cap=sim.cache_size, refit_every_substep=True,
                          velocity=velocity, mu=sim.mu, edge_edge=True)
    ne = cloth.num_edges
    for frame in range(8):
        print(f"[ee gate] velocity={velocity} frame={frame}", flush=True)
        sim.forward(5)
        ref.forward(5)
        np.testing.assert_array_equal(cloth.positions(), ref.x)
        live = sim._cd["ee"]["live"].cpu().numpy()
        print(f"[ee gate] live edges: static set {live[0]}, dynamic set {live[1 + ne]} of {ne}", flush=True)
        if frame < 2:                             # falling freely, not yet at the sphere
            assert live[1 + ne] == 0, "every cloth edge is gated while the cloth falls"
        assert 0 < live[0] <= ne                  # the sphere's edges are long: the static set runs
        np.testing.assert_array_equal(sim.edge_pairs()["num"].cpu().numpy(),
                                      np.concatenate([ref.last_ee["pn_st"], ref.last_ee["pn_dy"]]))
    assert ref.last_ee["pn_dy"].sum() == 0
