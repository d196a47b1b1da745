"""GPU: candidate lists served by the per-frame candidate cache equal the
traversal's, bit for bit (content, order, counts, overflow flags), whether a
query is served from the cache or walked again: small and large query moves,
refit trees whose vertices moved, caches too small for a query, tree
displacement beyond the margin, the per-lane and the wavefront walk."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def crumpled(slx, n, seed=0):
    g = slx.scene.grid_cloth(n, size=1.0, height=0.5, jitter=5e-3, seed=seed)
    x = g.x.astype(np.float64)
    x[:, 1] += 0.1 * np.sin(7 * x[:, 0]) * np.cos(5 * x[:, 2])
    return slx.Mesh(x=x.astype(np.float32), faces=g.faces)


def plain(A, tree, q4, pad, cap):
    import torch
    nq = q4.shape[0]
    cand = torch.full((cap * nq,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(nq, dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = A.ptr
    A.call("slx_traverse", 1, nq, P(q4), float(np.float32(pad)), tree.n_leaves, None, P(tree.sorted_faces),
           P(tree.nlo4), P(tree.nhi4), P(tree.quad4), None, None, None, cap, P(cand), P(cnt), P(st),
           A.stream_handle())
    return cand.view(cap, nq).cpu().numpy(), cnt.cpu().numpy(), st.item()


def cached(A, bp, tree, q4, pad, cap, cache, build, disp):
    import torch
    nq = q4.shape[0]
    cand = torch.full((cap * nq,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(nq, dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = A.ptr
    cs = cache.struct()
    A.call("slx_traverse_cached", 1 if build else 2, nq, P(q4), float(np.float32(pad)), tree.n_leaves,
           P(tree.nlo4), P(tree.nhi4), P(tree.quad4), A.C.byref(cs), P(disp), cap, P(cand), P(cnt), P(st), None,
           A.stream_handle())
    return cand.view(cap, nq).cpu().numpy(), cnt.cpu().numpy(), st.item()


def same(a, b):
    ca, na, sa = a
    cb, nb, sb = b
    np.testing.assert_array_equal(na, nb)
    assert sa == sb
    for q in range(len(na)):
        np.testing.assert_array_equal(ca[:na[q], q], cb[:nb[q], q])


@pytest.mark.parametrize("ccap,cap", [(80, 40), (6, 40), (80, 5)])
def test_cache_lists_equal_traversal(slx, rng, ccap, cap):
    import torch
    A = importlib.import_module("starlab-physics_amd._abi")
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    P = A.ptr
    cloth = crumpled(slx, 48)
    nv = cloth.num_verts
    pad, margin = np.float32(0.01), np.float32(0.01)
    tree = bp.DeviceTree(cloth, "tri")
    tree.build(float(pad))
    tree.enable_snapshot()
    x0 = cloth.positions()
    A.call("slx_tree_snapshot", nv, P(cloth.x4), P(tree.x_snap), P(tree.disp), A.stream_handle())
    cache = bp.CandidateCache(nv, ccap, float(margin), torch.device("cuda"))
    y0 = x0 + rng.uniform(-3e-3, 3e-3, x0.shape).astype(np.float32)
    q = torch.zeros((nv, 4), dtype=torch.float32, device="cuda")
    q[:, :3] = T(y0)
    print(f"[cache] ccap={ccap} cap={cap} build", flush=True)
    same(cached(A, bp, tree, q, pad, cap, cache, True, None), plain(A, tree, q, pad, cap))
    if ccap >= 80:
        assert cache.redone() == 0
    # the tree's vertices move (refit) and the queries move, some beyond the margin
    for step, (dx, big) in enumerate([(1e-3, 0.2), (3e-3, 0.3), (2e-2, 0.0)]):
        x1 = x0 + rng.uniform(-dx, dx, x0.shape).astype(np.float32)
        cloth.x4[:, :3] = T(x1)
        tree.refit(float(pad), cloth.x4)
        A.call("slx_tree_displacement", nv, P(cloth.x4), P(tree.x_snap), P(tree.disp), A.stream_handle())
        y1 = y0 + rng.uniform(-2e-3, 2e-3, y0.shape).astype(np.float32)
        far = rng.random(nv) < big
        y1[far] += rng.uniform(-2 * margin, 2 * margin, (far.sum(), 3)).astype(np.float32)
        q[:, :3] = T(y1)
        print(f"[cache] step {step} dx={dx}", flush=True)
        got = cached(A, bp, tree, q, pad, cap, cache, False, tree.disp)
        same(got, plain(A, tree, q, pad, cap))
        r = cache.redone()
        if dx >= 2e-2:
            assert r == nv                      # the tree moved more than the margin: every query walks
        elif ccap >= 80:
            assert 0 < r < nv                   # both paths taken
        assert np.asarray(got[1]).sum() > 0


def test_cache_wavefront_walk_of_a_coarse_tree(slx, rng):
    """A coarse tree (far fewer leaves than queries) is walked by the wavefront
    kernel; its cache build and filter give the same lists."""
    import torch
    A = importlib.import_module("starlab-physics_amd._abi")
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    ico = slx.scene.icosphere(1, 0.3, center=(0.0, 0.45, 0.0))
    obst = slx.Mesh(x=ico.x, faces=ico.faces, is_static=True)
    tree = bp.DeviceTree(obst, "tri")
    pad, margin = np.float32(0.02), np.float32(0.02)
    tree.build(float(pad))
    assert tree.n_leaves * 8 <= 2500
    n = 2500
    q = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    base = rng.uniform(-0.4, 0.4, (n, 3)).astype(np.float32) + np.float32([0, 0.45, 0])
    q[:, :3] = T(base)
    cache = bp.CandidateCache(n, 64, float(margin), torch.device("cuda"))
    same(cached(A, bp, tree, q, pad, 40, cache, True, None), plain(A, tree, q, pad, 40))
    moved = base + rng.uniform(-0.03, 0.03, base.shape).astype(np.float32)
    q[:, :3] = T(moved)
    got = cached(A, bp, tree, q, pad, 40, cache, False, None)
    same(got, plain(A, tree, q, pad, 40))
    assert 0 < cache.redone() < n
    assert got[1].sum() > 100


@pytest.mark.parametrize("depth", [16, 24, 32, 64])
def test_two_tier_walk_equals_traversal(slx, rng, depth):
    """The first-tier walk with a shallow stack, plus the full walk of the queries
    it could not serve, gives the traversal's lists, whatever the depth (large
    boxes overflow the shallow stacks: a 16-entry stack overflows for a few
    hundred queries at padding 0.1 on this cloth)."""
    import torch
    A = importlib.import_module("starlab-physics_amd._abi")
    bp = importlib.import_module("starlab-physics_amd.broadphase")
    P = A.ptr
    cloth = crumpled(slx, 48, seed=3)
    nv = cloth.num_verts
    tree = bp.DeviceTree(cloth, "tri")
    pad = np.float32(0.1)
    tree.build(float(pad))
    q = torch.zeros((nv, 4), dtype=torch.float32, device="cuda")
    q[:, :3] = T(cloth.positions() + rng.uniform(-0.02, 0.02, (nv, 3)).astype(np.float32))
    walk_only = bp.CandidateCache(nv, 0, 0.0, torch.device("cuda"))
    print(f"[two-tier] depth={depth}", flush=True)
    try:
        A.call("slx_set_walk_stack", depth)
        for cap in (40, 200):
            ref = plain(A, tree, q, pad, cap)
            nq = nv
            cand = torch.full((cap * nq,), -1, dtype=torch.int32, device="cuda")
            cnt = torch.zeros(nq, dtype=torch.int32, device="cuda")
            st = torch.zeros(1, dtype=torch.int32, device="cuda")
            cs = walk_only.struct()
            A.call("slx_traverse_cached", 0, nq, P(q), float(pad), tree.n_leaves, P(tree.nlo4), P(tree.nhi4),
                   P(tree.quad4), A.C.byref(cs), None, cap, P(cand), P(cnt), P(st), None, A.stream_handle())
            same((cand.view(cap, nq).cpu().numpy(), cnt.cpu().numpy(), st.item()), ref)
            if depth == 16:
                assert walk_only.redone() > 0          # the shallow stack overflowed somewhere
            cache = bp.CandidateCache(nv, 256, 0.005, torch.device("cuda"))
            same(cached(A, bp, tree, q, pad, cap, cache, True, None), ref)
    finally:
        A.call("slx_set_walk_stack", 24)   # the library default
