"""Pins of the oracle's narrow phase (distance.py, collision_constraints_x.py):

* d_type_PT region = Voronoi region of the closest point on the triangle
  (FP64 closest-point-on-triangle, Ericson, Real-Time Collision Detection 5.1.5),
  for points away from region boundaries;
* squared distance of that region's primitive = squared closest-point distance;
* gradients = central differences (FP64) and sum to zero;
* one contact in closed form (vertex above a static triangle).
"""
import numpy as np
import pytest

REGION = {"A": 0, "B": 1, "C": 2, "AB": 3, "BC": 4, "CA": 5, "F": 6}


def closest_region(p, a, b, c):
    """Ericson's ClosestPtPointTriangle, returning (region, closest point)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return "A", a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return "B", b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return "AB", a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return "C", c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return "CA", a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return "BC", b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return "F", a + ab * (vb * denom) + ac * (vc * denom)


def random_cases(rng, n):
    for _ in range(n):
        a, b, c = rng.normal(0, 1, (3, 3))
        p = rng.normal(0, 1.5, 3)
        yield p, a, b, c


def robust(p, a, b, c, eps=1e-3):
    """Region is stable under small perturbations of p (away from boundaries)."""
    r = closest_region(p, a, b, c)[0]
    for d in np.eye(3):
        for s in (eps, -eps):
            if closest_region(p + s * d, a, b, c)[0] != r:
                return False
    return True


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_region_matches_closest_point(orc, rng, dtype):
    checked = 0
    for p, a, b, c in random_cases(rng, 3000):
        if not robust(p, a, b, c):
            continue
        r, cp = closest_region(p, a, b, c)
        got = orc.d_type_pt(p, a, b, c, dtype)
        assert got == REGION[r], (r, got)
        d, _ = orc.dist(got, p, a, b, c, dtype)
        exact = float(np.sum((p - cp) ** 2))
        assert abs(d - exact) <= (1e-9 if dtype == np.float64 else 2e-4) * max(1.0, exact)
        checked += 1
    assert checked > 2000


def test_gradients_central_differences(orc, rng):
    h = 1e-6
    for p, a, b, c in random_cases(rng, 300):
        for dt_ in range(7):
            pts = [p.copy(), a.copy(), b.copy(), c.copy()]
            d0, g = orc.dist(dt_, *pts, np.float64)
            num = np.zeros((4, 3))
            for s in range(4):
                for k in range(3):
                    q1 = [x.copy() for x in pts]
                    q2 = [x.copy() for x in pts]
                    q1[s][k] += h
                    q2[s][k] -= h
                    num[s, k] = (orc.dist(dt_, *q1, np.float64)[0] - orc.dist(dt_, *q2, np.float64)[0]) / (2 * h)
            scale = max(1.0, np.abs(num).max())
            np.testing.assert_allclose(g, num, atol=2e-5 * scale, rtol=1e-5)
            np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-9 * scale)


def test_gradients_against_autograd(orc, rng):
    import torch

    def d_pp(p, a):
        return ((p - a) ** 2).sum()

    def d_pe(p, a, b):
        c = torch.linalg.cross(a - p, b - p)
        return (c ** 2).sum() / ((b - a) ** 2).sum()

    def d_pt(p, a, b, c):
        n = torch.linalg.cross(b - a, c - a)
        return ((p - a) @ n) ** 2 / (n @ n)
    for p, a, b, c in random_cases(rng, 50):
        T = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (p, a, b, c)]
        cases = {0: d_pp(T[0], T[1]), 4: d_pe(T[0], T[2], T[3]), 6: d_pt(*T)}
        for dt_, val in cases.items():
            gr = torch.autograd.grad(val, T, allow_unused=True)
            gr = np.stack([np.zeros(3) if x is None else x.numpy() for x in gr])
            d, g = orc.dist(dt_, p, a, b, c, np.float64)
            assert abs(d - val.item()) < 1e-10 * max(1, abs(d))
            np.testing.assert_allclose(g, gr, rtol=1e-9, atol=1e-10)


def test_vertex_above_static_triangle_closed_form(orc):
    """__vt_st (collision_constraints_x.py:5-69) for a vertex at height h above a
    static triangle: d = h^2, g0 = 2h n, ld = k (dhat - h^2) / (k w 4 h^2 + 1),
    y += w ld g0."""
    h, w, k, dhat = 0.005, 3.0, 1e8, 1e-4
    faces_dy = np.array([[1, 2, 3]], np.int32)     # unused by vt_st
    faces_st = np.array([[0, 1, 2]], np.int32)
    xs = np.array([[-1, 0, -1], [1, 0, -1], [0, 0, 1]], np.float64)
    y = np.array([[0.0, h, 0.0]])
    mesh = dict(faces=np.zeros((1, 3), np.int32))
    cand = (np.array([[0]], np.int32), np.array([1], np.int32))
    fixed = np.ones(1)
    m_inv = np.array([w])
    out = orc.collision_pass(mesh, faces_st, y, xs, fixed, m_inv, dhat, k, 40, cand, None, None)
    n = np.array([0.0, 1.0, 0.0])
    g0 = 2 * h * n
    ld = k * (dhat - h * h) / (k * w * (g0 @ g0) + 1)
    np.testing.assert_allclose(y[0], [0.0, h, 0.0] + w * ld * g0, rtol=1e-12)
    assert out["nc"][0] == 1 and out["pair_num"][0][0] == 1
    # outside dhat: untouched, no pair recorded
    y2 = np.array([[0.0, 0.5, 0.0]])
    out = orc.collision_pass(mesh, faces_st, y2, xs, fixed, m_inv, dhat, k, 40, cand, None, None)
    assert y2[0, 1] == 0.5 and out["pair_num"][0][0] == 0


def test_self_contact_skips_incident_faces_and_is_hard(orc):
    """__vt_dy: faces containing the vertex are skipped (XPBD.py:258); other pairs
    are recorded whatever d is; deltas use ld = (dhat - d)/(sum f w |g|^2 + 1e-4)."""
    faces = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    y = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0.2, 0.003, 0.2], [3, 1, 3], [4, 1, 3]], np.float64)
    fixed = np.ones(6)
    m_inv = np.ones(6)
    mesh = dict(faces=faces)
    # vertex 3 against face 0 (close), vertex 0 against its own face 0 (skipped)
    cand = np.full((6, 2), -1, np.int32)
    cnt = np.zeros(6, np.int32)
    cand[3, 0], cnt[3] = 0, 1
    cand[0, 0], cnt[0] = 0, 1
    y0 = y.copy()
    out = orc.collision_pass(mesh, None, y, None, fixed, m_inv, 1e-4, 1e8, 2, None, (cand, cnt), None)
    assert out["pair_num"][1][0] == 0 and out["pair_num"][1][3] == 1
    d = 0.003 ** 2
    g = 2 * 0.003
    schur = g * g + 3 * (g / 3) ** 2 * 0 + 1e-4
    assert y[3, 1] > y0[3, 1] and y[0, 1] < y0[0, 1]       # pushed apart along the normal
    assert np.isclose(out["nc"][3], 1) and np.isclose(out["nc"][0], 1)
    # total momentum (unit masses) is conserved by the symmetric hard constraint
    np.testing.assert_allclose((y - y0).sum(axis=0), 0.0, atol=1e-12)
    assert d < 1e-4 and schur > 0
