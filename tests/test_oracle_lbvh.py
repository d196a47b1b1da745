"""Pins of the oracle's LBVH (framework/collision/lbvh_cell.py) by closed forms
and brute force:

* expand_bits / Morton: bit-interleave closed form over all 1024 values per axis;
* cell codes: code(x0,y0,z0) = E(32x0+16) | E(32y0+16)<<1 | E(32z0+16)<<2;
* radix sort: numpy stable argsort;
* Karras: the unique binary radix tree built top-down by splitting at the
  highest differing bit (index tie-break for equal codes);
* cell build: stable bucketing, leaf boxes and refit by brute-force min/max;
* traversal: candidate sets = brute-force box overlap; order = recursive DFS
  visiting child_b before child_a.
"""
import numpy as np
import pytest

NC = 32768


def interleave(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_expand_bits_closed_form(orc):
    for v in range(1024):
        assert orc.expand_bits(v) == interleave(v)
    assert orc.expand_bits(1023) == 0x09249249


def test_morton_known_answers(orc):
    assert orc.morton3d(1.0, 1.0, 1.0) == 0x3FFFFFFF
    assert orc.morton3d(0.0, 0.0, 0.0) == 0
    assert orc.morton3d(1.0, 0.0, 0.0) == interleave(1023)          # x in the lowest bit
    assert orc.morton3d(0.0, 1.0, 0.0) == interleave(1023) << 1
    assert orc.morton3d(-3.0, 7.0, float("nan")) == interleave(1023) << 1   # clamped


def test_cell_codes_closed_form(orc):
    codes, centers, cs = orc.cell_morton(np.zeros(3, np.float32), np.full(3, 1e3, np.float32))
    i = np.arange(NC)
    x0, y0, z0 = i // 1024, (i % 1024) // 32, i % 32
    E = np.array([interleave(v) for v in range(1024)], np.int64)
    expect = E[32 * x0 + 16] | (E[32 * y0 + 16] << 1) | (E[32 * z0 + 16] << 2)
    np.testing.assert_array_equal(codes, expect.astype(np.int32))
    assert codes[0] == 0x7000 and codes[1] == 0x27000 and codes[32] == 0x17000 and codes[1024] == 0xF000
    assert codes.min() == 0x7000 and codes.max() == 0x3FFFF000
    assert len(np.unique(codes)) == NC


@pytest.mark.parametrize("bits,passes", [(6, 5), (8, 4)])
def test_radix_sort_is_stable_argsort(orc, rng, bits, passes):
    keys = rng.integers(0, 1 << 30, 5000, dtype=np.int64).astype(np.int32)
    keys[::7] = keys[3]                       # many duplicates
    k, ids = orc.radix_sort(keys, np.arange(len(keys), dtype=np.int32), bits, passes)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(ids, order)
    np.testing.assert_array_equal(k, keys[order])


def topdown_tree(codes, wide):
    """Binary radix tree by recursive splitting (distinct keys, or (code, index)
    keys when wide).  Returns {internal node index: (left, right)} with leaves as
    ('L', k) and internal nodes identified by their covered range."""
    n = len(codes)
    keys = [((int(c) & 0xFFFFFFFF) << 32 | i) if wide else (int(c) & 0xFFFFFFFF) for i, c in enumerate(codes)]
    nbits = 64 if wide else 32
    out = {}

    def rec(lo, hi):
        if lo == hi:
            return ("L", lo)
        x = keys[lo] ^ keys[hi]
        bit = x.bit_length() - 1
        split = lo
        while split + 1 <= hi and not ((keys[split + 1] >> bit) & 1):
            split += 1
        a, b = rec(lo, split), rec(split + 1, hi)
        out[(lo, hi)] = (a, b)
        return ("N", lo, hi)
    if n > 1:
        rec(0, n - 1)
    assert nbits
    return out


def karras_as_ranges(t, n):
    """Convert orc.karras output to the same {range: (child, child)} form."""
    ca, cb, rl, rr = t["child_a"], t["child_b"], t["range_l"], t["range_r"]
    def node(k):
        if k >= n - 1:
            return ("L", k - (n - 1))
        return ("N", int(rl[k]), int(rr[k]))
    return {(int(rl[i]), int(rr[i])): (node(ca[i]), node(cb[i])) for i in range(n - 1)}


@pytest.mark.parametrize("n,dups", [(2, False), (3, False), (37, False), (500, False), (500, True), (64, True)])
def test_karras_equals_topdown_radix_tree(orc, rng, n, dups):
    if dups:
        codes = np.sort(rng.integers(0, 40, n)).astype(np.int32)
    else:
        codes = np.sort(rng.choice(1 << 30, n, replace=False)).astype(np.int32)
    t = orc.karras(codes, wide=dups)
    assert karras_as_ranges(t, n) == topdown_tree(codes, dups)
    # every node except the root has a parent, root's parent is -1
    assert t["parent"][0] == -1
    assert np.all(t["parent"][1:] >= 0)


def test_cell_topology_is_radix_tree_of_cell_codes(orc):
    topo = orc.cell_topology()
    assert karras_as_ranges(dict(child_a=topo.child_a, child_b=topo.child_b, range_l=topo.range_l,
                                 range_r=topo.range_r), NC) == topdown_tree(topo.sorted_codes, False)


def make_faces(slx, n=20, seed=0):
    g = slx.scene.grid_cloth(n, size=1.0, height=0.5, jitter=5e-3, seed=seed)
    # crumple it so boxes overlap in 3-D
    x = g.x.copy()
    x[:, 1] += 0.1 * np.sin(7 * x[:, 0]) * np.cos(5 * x[:, 2])
    return x.astype(np.float32), g.faces


def test_cell_build_brute_force(slx, orc):
    x, faces = make_faces(slx)
    pad = 0.01
    t = orc.CellBVH(faces, x, pad)
    assert t.bad == 0
    # stable bucketing: ascending face id inside each cell
    np.testing.assert_array_equal(t.sorted_faces[:len(faces)], np.argsort(t.face_cell, kind="stable"))
    counts = np.bincount(t.face_cell, minlength=NC)
    np.testing.assert_array_equal(t.prefix, np.cumsum(counts))
    # leaf boxes
    lo = t.lo
    cs = (t.hi - t.lo) / np.float32(32)
    for cell in np.nonzero(counts)[0][:50]:
        node = cell + NC - 1
        fs = np.nonzero(t.face_cell == cell)[0]
        i, j, k = cell // 1024, (cell % 1024) // 32, cell % 32
        c = ((cs * np.array([i, j, k], np.float32)) + (np.float32(0.5) * cs)) + lo
        mn = np.minimum(c, t.fmin[fs].min(axis=0))
        mx = np.maximum(c, t.fmax[fs].max(axis=0))
        np.testing.assert_array_equal(t.nmin[node], mn)
        np.testing.assert_array_equal(t.nmax[node], mx)
        assert t.num[node] == len(fs)
    # root covers all non-empty leaves and counts every face
    assert t.num[0] == len(faces)
    ne = np.nonzero(counts)[0] + NC - 1
    np.testing.assert_array_equal(t.nmin[0], t.nmin[ne].min(axis=0))
    np.testing.assert_array_equal(t.nmax[0], t.nmax[ne].max(axis=0))


def brute_candidates(qmin, qmax, fmin, fmax):
    ov = np.all((qmin[:, None, :] <= fmax[None]) & (qmax[:, None, :] >= fmin[None]), axis=2)
    return [set(np.nonzero(r)[0].tolist()) for r in ov]


def dfs_order(t, a, b, cell):
    """Recursive restatement of the reference traversal order."""
    out = []
    nl = t.n_leaves

    def ov(lo, hi):
        return np.all(a <= hi) and np.all(b >= lo)

    def visit(i):
        if cell and t.num[i] <= 0:
            return
        if not ov(t.nmin[i], t.nmax[i]):
            return
        if i >= nl - 1:
            leaf = i - (nl - 1)
            rng_ = range(t.leaf_l[leaf], t.leaf_r[leaf] + 1) if cell else [leaf]
            for jj in rng_:
                f = t.sorted_faces[jj]
                if ov(t.fmin[f], t.fmax[f]):
                    out.append(int(f))
        else:
            visit(t.child_b[i])      # pushed last, popped first
            visit(t.child_a[i])
    visit(0)
    return out


@pytest.mark.parametrize("mode", ["cell", "tri"])
def test_traversal_sets_and_order(slx, orc, rng, mode):
    x, faces = make_faces(slx, n=16)
    pad = 0.02
    T = orc.CellBVH if mode == "cell" else orc.TriBVH
    t = T(faces, x, pad)
    q = x[rng.choice(len(x), 60, replace=False)] + rng.normal(0, 0.01, (60, 3)).astype(np.float32)
    qmin, qmax = (q - np.float32(pad)).astype(np.float32), (q + np.float32(pad)).astype(np.float32)
    cand, cnt, fl = t.query(qmin, qmax, 200)
    assert fl == 0
    sets = brute_candidates(qmin, qmax, t.fmin, t.fmax)
    for i in range(len(q)):
        got = cand[i, :cnt[i]].tolist()
        assert set(got) == sets[i]
        assert got == dfs_order(t, qmin[i], qmax[i], mode == "cell")


def test_traversal_cap_and_overflow(slx, orc):
    x, faces = make_faces(slx, n=12)
    t = orc.TriBVH(faces, x, 0.2)
    q = x[:5]
    cand, cnt, fl = t.query(q - 0.2, q + 0.2, 3)
    assert np.all(cnt == 3) and fl == 5


def test_tri_tree_refit_matches_rebuild_boxes(slx, orc):
    x, faces = make_faces(slx, n=14)
    t = orc.TriBVH(faces, x, 0.01)
    x2 = (x + np.float32(0.003)).astype(np.float32)
    t.refit(x2)
    # internal boxes = min/max over the leaves below (brute force on the root)
    np.testing.assert_array_equal(t.nmin[0], t.fmin.min(axis=0))
    np.testing.assert_array_equal(t.nmax[0], t.fmax.max(axis=0))


def test_query_cell_filter_skips_out_of_grid(slx, orc):
    """lbvh_cell.py:665: a query whose centre cell id is > 32767 is skipped."""
    x, faces = make_faces(slx, n=10)
    t = orc.CellBVH(faces, x, 0.01)
    far = np.array([[100.0, 0.5, 0.0]], np.float32)
    cand, cnt, _ = t.query(far - 0.01, far + 0.01, 10)
    assert cnt[0] == 0


def test_tri_codes_normalise_by_the_largest_extent(slx, orc):
    """Triangle-tree codes: face-box centres normalised by one scale, the mesh box's
    largest extent e, on every axis (a flat sheet keeps its bits for its width).
    Known answers: the face at the far x end of a sheet 4x wider than tall codes
    x = 1 and y within the lower quarter; and the codes equal an independent
    numpy restatement for every face."""
    x, faces = make_faces(slx, n=9)
    x = x.astype(np.float64)
    x[:, 1] = 0.25 * x[:, 0] * x[:, 0]          # a gentle slope: y spans less than x and z
    x = x.astype(np.float32)
    pad = np.float32(0.01)
    t = orc.TriBVH(faces, x, pad)
    fmin, fmax = t.fmin.astype(np.float32), t.fmax.astype(np.float32)
    lo, hi = t.lo.astype(np.float32), t.hi.astype(np.float32)
    e = np.float32(max(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]))
    top = (lo + e).astype(np.float32)
    c = (np.float32(0.5) * (fmin + fmax)).astype(np.float32)
    u = np.clip((c - lo) / (top - lo), 0, 1).astype(np.float32)
    want = np.array([orc.morton3d(*map(float, r)) for r in u], np.int64)
    np.testing.assert_array_equal(t.codes.astype(np.int64), want)
    # per-axis normalisation would have coded y over its own extent: different codes
    u_axis = np.clip((c - lo) / (hi - lo), 0, 1).astype(np.float32)
    per_axis = np.array([orc.morton3d(*map(float, r)) for r in u_axis], np.int64)
    assert (per_axis != want).any()
    # y bits stay below a quarter of the cube (the sheet is at most ~1/4 as tall as wide)
    assert float(u[:, 1].max()) <= 0.3
