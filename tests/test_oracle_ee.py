"""Pins of the oracle's edge-edge constraints (framework/collision/distance.py
49-124, framework/physics/collision_constraints_x.py:312-507,
collision_constraints_v.py:573-878, the loops of XPBD.py:266-281 / 313-330).

* d_EE against an independent computation: the squared distance of the two
  lines from the minimiser of |x0 + s u - x2 - t v|^2 (a 2x2 least-squares
  solve); its gradient against central differences in float64.
* d_type_EE against brute force: the region (vertex pair, vertex-edge, both
  interiors) of the exact segment-segment minimum over all nine candidate
  sub-problems, where the reference's rule is the textbook one; where its extra
  condition (u x v).w < 1e-4 with w = x0 - x2 applies, the interior type 8
  never occurs.
* The position constraints: for two perpendicular edges crossing at their
  midpoints a height h apart, d = h^2, g = (0, h, 0) on the upper edge's
  points and (0, -h, 0) on the lower one's; ld = k (dhat - d) / (k schur + 1)
  and dx_k = w_k ld g_k, with schur over the moving points only (two for the
  static set, four for the dynamic set); nc counts them.  Parallel edges, edges
  sharing a vertex (dynamic set), distances >= dhat and endpoint regions act on
  nothing and record nothing.
* The velocity pass: __ee_st's Cv leaves out g2 (the static edge's first
  point's velocity has no effect, the second's has), __ee_dy cancels the
  approaching normal velocity over the four points; both only when closing.
"""
import numpy as np
import pytest


def brute_type(x0, x1, x2, x3):
    """Region of the minimum of |x0 + s (x1-x0) - x2 - t (x3-x2)|^2 on [0,1]^2,
    as (type, gap to the second-best region)."""
    u, v, w = x1 - x0, x3 - x2, x0 - x2

    def f(s, t):
        r = w + s * u - t * v
        return r @ r

    def clamp01(a):
        return min(max(a, 0.0), 1.0)

    cands = {}
    A = np.array([[u @ u, -(u @ v)], [-(u @ v), v @ v]])
    st = np.linalg.solve(A, np.array([-(u @ w), v @ w]))
    if 0 < st[0] < 1 and 0 < st[1] < 1:
        cands[8] = f(*st)
    cands[0] = f(0, 0)
    cands[1] = f(0, 1)
    cands[3] = f(1, 0)
    cands[4] = f(1, 1)
    for typ, s in ((2, 0.0), (5, 1.0)):                 # s fixed, t interior
        t = ((w + s * u) @ v) / (v @ v)
        if 0 < t < 1:
            cands[typ] = f(s, t)
    for typ, t in ((6, 0.0), (7, 1.0)):                 # t fixed, s interior
        s = -((w - t * v) @ u) / (u @ u)
        if 0 < s < 1:
            cands[typ] = f(s, t)
    order = sorted(cands.items(), key=lambda kv: kv[1])
    gap = order[1][1] - order[0][1] if len(order) > 1 else np.inf
    return order[0][0], gap


def test_d_ee_is_the_squared_line_distance(orc, rng):
    for _ in range(200):
        x = rng.normal(size=(4, 3))
        d, _ = orc.dist_ee(*x, dtype=np.float64)
        u, v = x[1] - x[0], x[3] - x[2]
        M = np.stack([u, -v], axis=1)
        st, *_ = np.linalg.lstsq(M, x[2] - x[0], rcond=None)
        r = x[0] + st[0] * u - x[2] - st[1] * v
        assert np.isclose(d, r @ r, rtol=1e-9, atol=1e-14)


def test_d_ee_gradient_central_differences(orc, rng):
    h = 1e-6
    for _ in range(100):
        x = rng.normal(size=(4, 3))
        _, g = orc.dist_ee(*x, dtype=np.float64)
        num = np.zeros((4, 3))
        for k in range(4):
            for c in range(3):
                xp, xm = x.copy(), x.copy()
                xp[k, c] += h
                xm[k, c] -= h
                num[k, c] = (orc.dist_ee(*xp, dtype=np.float64)[0] - orc.dist_ee(*xm, dtype=np.float64)[0]) / (2 * h)
        assert np.allclose(g, num, rtol=1e-5, atol=1e-7), (g, num)
        # translation invariance: the four gradients sum to zero
        assert np.allclose(g.sum(axis=0), 0.0, atol=1e-10)


def test_d_ee_float32_agrees_with_float64(orc, rng):
    for _ in range(100):
        x = rng.normal(size=(4, 3)).astype(np.float32)
        d32, g32 = orc.dist_ee(*x, dtype=np.float32)
        d64, g64 = orc.dist_ee(*x.astype(np.float64), dtype=np.float64)
        assert np.isclose(d32, d64, rtol=1e-3, atol=1e-6)
        assert np.allclose(g32, g64, rtol=1e-2, atol=1e-4 * np.abs(g64).max())


def test_d_type_ee_brute_force(orc, rng):
    seen = set()
    quirk = 0
    for _ in range(4000):
        x = rng.normal(size=(4, 3))
        x[2:] += rng.normal(scale=0.5, size=3)
        typ = orc.d_type_ee(*x, dtype=np.float64)
        u, v, w = x[1] - x[0], x[3] - x[2], x[0] - x[2]
        triple = np.cross(u, v) @ w
        bt, gap = brute_type(*x)
        if gap < 1e-9:
            continue
        if triple >= 1e-4:
            assert typ == bt, (x, typ, bt)
            seen.add(typ)
        else:
            assert typ != 8
            if bt != 8:                      # the extra condition only touches interior minima
                assert typ == bt, (x, typ, bt)
                seen.add(typ)
            else:
                quirk += 1
    assert seen == set(range(9))
    assert quirk > 100


def test_d_type_ee_float32_matches_float64_away_from_boundaries(orc, rng):
    agree = 0
    for _ in range(500):
        x = rng.normal(size=(4, 3)).astype(np.float32)
        t32 = orc.d_type_ee(*x, dtype=np.float32)
        t64 = orc.d_type_ee(*x.astype(np.float64), dtype=np.float64)
        agree += t32 == t64
    assert agree >= 495


# ------------------------------------------------------------- position pass

DHAT = 1e-2


def cross_scene(h, mirrored=False):
    """Positions of two triangles: (0,1,2) with edge (0,1) along +x at height h,
    (3,4,5) with edge (3,4) along -z at height 0, crossing under the first at
    the origin.  With that orientation (u x v).w = 4h > 0 for both edges' queries
    (u, v the edges' directions, w = x0 - x2), so the interior type is reached;
    mirrored=True runs the lower edge along +z instead, where the reference's
    extra condition forces the pair off the interior type."""
    z = -1.0 if mirrored else 1.0
    return np.array([[-1.0, h, 0.0], [1.0, h, 0.0], [0.0, h, 2.0],
                     [0.0, 0.0, z], [0.0, 0.0, -z], [-2.0, 0.0, 0.0]])


def topo(slx, faces):
    e = slx.topology.edges_from_faces(faces)
    return e, slx.topology.face_edges(faces, e)


def test_face_edges_topology(slx):
    faces = np.array([[0, 1, 2], [2, 1, 3]], np.int32)
    e, fe = topo(slx, faces)
    np.testing.assert_array_equal(e, [[0, 1], [0, 2], [1, 2], [1, 3], [2, 3]])
    np.testing.assert_array_equal(fe, [[0, 2, 1], [2, 3, 4]])


def ee_st_case(orc, slx, h, *, pcap=6, k=1e8, dhat=DHAT, mirrored=False):
    """Upper triangle dynamic (mesh_dy), lower triangle static."""
    y_all = cross_scene(h, mirrored)
    faces_dy = np.array([[0, 1, 2]], np.int32)
    faces_st = np.array([[0, 1, 2]], np.int32)
    y = y_all[:3].copy()
    xs = y_all[3:].copy()
    e_dy, fe_dy = topo(slx, faces_dy)
    e_st, fe_st = topo(slx, faces_st)
    ne = len(e_dy)
    cand = np.zeros((ne, 2), np.int32)
    cnt = np.ones(ne, np.int32)
    w = np.array([1.0, 2.0, 3.0])
    fixed = np.ones(3)
    ee = dict(edges_dy=e_dy, edges_st=e_st, fe_dy=fe_dy, fe_st=fe_st, st=(cand, cnt), dy=None, pcap=pcap)
    out = orc.collision_pass(dict(faces=faces_dy), faces_st, y, xs, fixed, w, dhat, k, 2, None, None, None,
                             record=True, ee=ee)
    return y_all, y, out, w, e_dy, e_st


def test_ee_st_crossing_closed_form(orc, slx):
    h, k = 0.05, 1e8
    y0, y, out, w, e_dy, e_st = ee_st_case(orc, slx, h, k=k)
    E = out["ee"]
    # dynamic edge 0 = (0, 1) along x; static edge (0, 1) = (3, 4) along z
    assert list(e_dy[0]) == [0, 1] and list(e_st[0]) == [0, 1]
    assert E["pn_st"][0] == 1                        # only the crossing pair contacts
    assert E["pn_st"][1:].sum() == 0
    assert E["pi_st"][0, 0, 0] == 0 and E["pi_st"][0, 0, 1] == 8
    g = E["pg_st"][0, 0]
    np.testing.assert_allclose(g, [[0, h, 0], [0, h, 0], [0, -h, 0], [0, -h, 0]], atol=1e-12)
    d = h * h
    schur = (w[0] + w[1]) * h * h
    assert np.isclose(E["ps_st"][0, 0], schur, rtol=1e-12)
    ld = k * (DHAT - d) / (k * schur + 1.0)
    # Jacobi update: y += dx / nc with nc = 1 on the two points
    np.testing.assert_allclose(y[0], y0[0] + w[0] * ld * np.array([0, h, 0]), rtol=1e-12)
    np.testing.assert_allclose(y[1], y0[1] + w[1] * ld * np.array([0, h, 0]), rtol=1e-12)
    np.testing.assert_array_equal(y[2], y0[2])
    np.testing.assert_array_equal(out["nc"], [1, 1, 0])
    # the same crossing with the static edge reversed: (u x v).w = -4h < 1e-4, so
    # the interior minimum (s = 1/2, sN = D/2 exactly) is moved to the s = 1 side,
    # type 5 (x1 against the static edge) instead of 8, and nothing acts
    y0m, ym, outm, *_ = ee_st_case(orc, slx, h, k=k, mirrored=True)
    assert outm["ee"]["pn_st"].sum() == 0
    assert orc.d_type_ee(y0m[0], y0m[1], y0m[3], y0m[4], dtype=np.float64) == 5
    np.testing.assert_array_equal(ym, y0m[:3])


def test_ee_st_out_of_range_and_parallel(orc, slx):
    h = 0.2                                          # d = 0.04 >= dhat
    y0, y, out, *_ = ee_st_case(orc, slx, h)
    assert out["ee"]["pn_st"].sum() == 0
    np.testing.assert_array_equal(y, y0[:3])
    # parallel edges: the static triangle rotated so its edge runs along x too
    faces = np.array([[0, 1, 2]], np.int32)
    e, fe = topo(slx, faces)
    y = np.array([[-1.0, 0.01, 0.0], [1.0, 0.01, 0.0], [0.0, 0.01, 2.0]])
    xs = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, -2.0]])
    ee = dict(edges_dy=e, edges_st=e, fe_dy=fe, fe_st=fe, st=(np.zeros((3, 1), np.int32), np.ones(3, np.int32)))
    out = orc.collision_pass(dict(faces=faces), faces, y.copy(), xs, np.ones(3), np.ones(3), DHAT, 1e8, 1, None,
                             None, None, record=True, ee=ee)
    assert out["ee"]["pn_st"][0] == 0                # x01 x x23 = 0: no constraint
    np.testing.assert_array_equal(out["nc"], [0, 0, 0])


def test_ee_dy_crossing_closed_form_and_shared_vertex(orc, slx):
    h, k = 0.05, 1e8
    y0 = cross_scene(h)
    faces = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    e, fe = topo(slx, faces)
    ne = len(e)
    y = y0.copy()
    w = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    fixed = np.array([1.0, 1.0, 1.0, 1.0, 0.0, 1.0])          # vertex 4 pinned
    # every edge sees both faces: the shared-vertex skip removes its own face's edges
    cand = np.tile(np.array([[0, 1]], np.int32), (ne, 1))
    cnt = np.full(ne, 2, np.int32)
    ee = dict(edges_dy=e, edges_st=e, fe_dy=fe, fe_st=fe, st=None, dy=(cand, cnt))
    out = orc.collision_pass(dict(faces=faces), None, y, None, fixed, w, DHAT, k, 2, None, None, None,
                             record=True, ee=ee)
    E = out["ee"]
    i01 = int(np.nonzero((e[:, 0] == 0) & (e[:, 1] == 1))[0][0])
    i34 = int(np.nonzero((e[:, 0] == 3) & (e[:, 1] == 4))[0][0])
    # the pair acts once from each side (edge 01's query and edge 34's query)
    assert E["pn_dy"][i01] == 1 and E["pn_dy"][i34] == 1
    assert E["pn_dy"].sum() == 2
    g01 = E["pg_dy"][i01, 0]
    np.testing.assert_allclose(g01, [[0, h, 0], [0, h, 0], [0, -h, 0], [0, -h, 0]], atol=1e-12)
    schur = (w[0] + w[1] + w[3] + 0.0 * w[4]) * h * h
    assert np.isclose(E["ps_dy"][i01, 0], schur, rtol=1e-12)
    ld = k * (DHAT - h * h) / (k * schur + 1.0)
    # both sides give the same deltas (the second query sees the pair mirrored)
    up = w[0] * ld * h
    np.testing.assert_allclose(y[0], y0[0] + [0, up, 0], rtol=1e-12)
    np.testing.assert_allclose(y[3], y0[3] - [0, w[3] * ld * h, 0], rtol=1e-12)
    np.testing.assert_array_equal(y[4], y0[4])                # pinned
    np.testing.assert_array_equal(out["nc"], [2, 2, 0, 2, 2, 0])


def test_ee_records_past_pcap_are_counted(orc, slx):
    """The reference keeps no bound on its edge pair fields; the oracle counts
    every contacting pair and stores the first pcap."""
    h = 0.05
    y0 = cross_scene(h)
    faces_dy = np.array([[0, 1, 2]], np.int32)
    faces_st = np.array([[0, 1, 2]], np.int32)
    e, fe = topo(slx, faces_dy)
    cand = np.zeros((3, 3), np.int32)                 # the same face three times
    cnt = np.full(3, 3, np.int32)
    ee = dict(edges_dy=e, edges_st=e, fe_dy=fe, fe_st=fe, st=(cand, cnt), pcap=2)
    out = orc.collision_pass(dict(faces=faces_dy), faces_st, y0[:3].copy(), y0[3:].copy(), np.ones(3), np.ones(3),
                             DHAT, 1e8, 3, None, None, None, record=True, ee=ee)
    assert out["ee"]["pn_st"][0] == 3
    np.testing.assert_array_equal(out["nc"], [3, 3, 0])


# ------------------------------------------------------------- velocity pass

def test_ee_st_velocity_leaves_out_g2(orc, slx):
    h = 0.05
    y0, y, out, w, e_dy, e_st = ee_st_case(orc, slx, h)
    E = out["ee"]
    mesh = dict(faces=np.array([[0, 1, 2]], np.int32))
    fst = np.array([[0, 1, 2]], np.int32)
    g = E["pg_st"][0, 0]
    schur = E["ps_st"][0, 0]

    def run(v_dyn, vs):
        v = np.array(v_dyn, float)
        orc.velocity_pass(mesh, fst, v, np.array(vs, float), np.ones(3), w, None, 0.8, None, 2, ee=E)
        return v

    # approaching (the upper edge moves down): Cv = g0.v0 + g1.v1 < 0
    v_in = [[0, -1.0, 0], [0, -0.5, 0], [0, 0, 0]]
    vs0 = np.zeros((3, 3))
    v = run(v_in, vs0)
    Cv = g[0] @ v_in[0] + g[1] @ v_in[1]
    ld = -Cv / schur
    np.testing.assert_allclose(v[0], np.array(v_in[0]) + w[0] * ld * g[0], rtol=1e-12)
    np.testing.assert_allclose(v[1], np.array(v_in[1]) + w[1] * ld * g[1], rtol=1e-12)
    # the static edge's first point (v2 = static vertex 0) does not enter Cv ...
    vs_a = vs0.copy()
    vs_a[0] = [0, 100.0, 0]
    np.testing.assert_array_equal(run(v_in, vs_a), v)
    # ... its second point (static vertex 1) does: moving up at 100 makes Cv more negative
    vs_b = vs0.copy()
    vs_b[1] = [0, 100.0, 0]
    v_b = run(v_in, vs_b)
    Cv_b = Cv + g[3] @ vs_b[1]
    np.testing.assert_allclose(v_b[0], np.array(v_in[0]) + w[0] * (-Cv_b / schur) * g[0], rtol=1e-12)
    # separating: nothing
    v_out = [[0, 1.0, 0], [0, 0.5, 0], [0, 0, 0]]
    np.testing.assert_array_equal(run(v_out, vs0), v_out)


def test_ee_dy_velocity_cancels_approach(orc, slx):
    h, k = 0.05, 1e8
    y0 = cross_scene(h)
    faces = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    e, fe = topo(slx, faces)
    ne = len(e)
    w = np.ones(6)
    cand = np.tile(np.array([[0, 1]], np.int32), (ne, 1))
    ee = dict(edges_dy=e, edges_st=e, fe_dy=fe, fe_st=fe, st=None, dy=(cand, np.full(ne, 2, np.int32)))
    out = orc.collision_pass(dict(faces=faces), None, y0.copy(), None, np.ones(6), w, DHAT, k, 2, None, None, None,
                             record=True, ee=ee)
    E = out["ee"]
    v_in = np.zeros((6, 3))
    v_in[[0, 1], 1] = -1.0                              # upper edge falls, lower edge rises
    v_in[[3, 4], 1] = 1.0
    v = v_in.copy()
    orc.velocity_pass(dict(faces=faces), None, v, None, np.ones(6), w, None, 0.8, None, 2, nvs=0, ee=E)
    # dvn = g . v = 4 (-h) = -4h; schur = 4 h^2; ld = -1/h; dv_k = -g_k ld = g_k / h,
    # twice (both queries), averaged over nc = 2: each point's normal velocity cancels
    np.testing.assert_allclose(v[[0, 1], 1], 0.0, atol=1e-12)
    np.testing.assert_allclose(v[[3, 4], 1], 0.0, atol=1e-12)
    np.testing.assert_array_equal(v[[2, 5]], v_in[[2, 5]])
