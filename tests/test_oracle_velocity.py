"""Pins of the oracle's velocity-level contact and friction pass
(framework/physics/collision_constraints_v.py:5-570, XPBD.py:283-311, 374-404)
and of the pair records the position pass leaves for it
(collision_constraints_x.py:60-68, 164-172, 301-309).

Each case is a single contact whose outcome has a closed form in terms of the
contact normal n (unit vector of the squared-distance gradient g0):

* normal part: the relative normal velocity is cancelled, dv_n = -(v.n) n for a
  vertex against a static primitive (w_k weights of the static point);
* friction: with vT the post-normal tangential (relative) velocity and s =
  w |vT|^2 / (w |vT|^2 + 1e-4), dv_t = -(s/2) vT; applied in full when
  mu |Cv| > |vT|^2 / 2 (static), else scaled by mu (kinetic);
* update_dv: v += f dv / nc with nc the number of deltas the vertex received.

The reference's quirks are pinned the same way: type 5 against a static edge
flips the friction direction (collision_constraints_v.py:109), type 3 takes the
inverse mass of the static mesh at the dynamic vertex's index (:64), __tv_st uses
g3 for t1's velocity in Cv (:144), __vt_dy type 1 scales t1's friction delta by
t0's inverse mass (:388).  A plausible mistake (a dropped factor, the "fixed"
formula instead of the reference's, a wrong sign) fails one of them.
"""
import numpy as np
import pytest

TRI = np.array([[-1.0, 0.0, -1.0], [1.0, 0.0, -1.0], [0.0, 0.0, 1.0]])
DHAT = 1e-4
EPS = float(np.float32(1e-4))   # the regulariser is a float32 literal in the reference's float32 kernels


def unit(a):
    return a / np.linalg.norm(a)


def friction_scale(w, vt):
    dd = vt @ vt
    return (w * dd) / (w * dd + EPS)


def vt_st_case(orc, p, v0, *, w=2.0, mu=0.8, vs=None, m_inv_st=None, dyn_index=0):
    """One dynamic vertex (index dyn_index) against the static triangle TRI."""
    nv = dyn_index + 1
    faces_st = np.array([[0, 1, 2]], np.int32)
    y = np.zeros((nv, 3))
    y[dyn_index] = p
    mesh = dict(faces=np.zeros((1, 3), np.int32))
    cand = np.full((nv, 4), -1, np.int32)
    cnt = np.zeros(nv, np.int32)
    cand[dyn_index, 0], cnt[dyn_index] = 0, 1
    fixed = np.ones(nv)
    m_inv = np.full(nv, w)
    out = orc.collision_pass(mesh, faces_st, y, TRI.copy(), fixed, m_inv, DHAT, 1e8, 4, (cand, cnt), None, None,
                             record=True)
    rec = out["pairs"][0]
    v = np.zeros((nv, 3))
    v[dyn_index] = v0
    orc.velocity_pass(mesh, faces_st, v, vs, fixed, m_inv, m_inv_st, mu, out["pairs"], 4)
    return rec, v[dyn_index]


def test_vt_st_face_record_and_closed_form(orc):
    h, w = 0.005, 2.0
    p = np.array([0.0, h, 0.0])
    v0 = np.array([0.3, -0.2, 0.1])
    for mu, static in ((0.8, False), (100.0, True)):
        rec, v = vt_st_case(orc, p, v0, w=w, mu=mu)
        # the record: face 0, type 6, g0 = 2h n, g_k = -b_k g0 (barycentrics .25, .25, .5),
        # schur = w |g0|^2 (collision_constraints_x.py:56-68)
        assert rec["num"][0] == 1 and rec["fid"][0, 0] == 0 and rec["dtype"][0, 0] == 6
        g = rec["g"][0, 0]
        np.testing.assert_allclose(g[0], [0, 2 * h, 0], rtol=1e-12)
        np.testing.assert_allclose(g[1:], -np.outer([0.25, 0.25, 0.5], g[0]), rtol=1e-10, atol=1e-16)
        assert np.isclose(rec["schur"][0, 0], w * 4 * h * h, rtol=1e-12, atol=0)
        n = np.array([0.0, 1.0, 0.0])
        dv_n = -(v0 @ n) * n
        vt = v0 + dv_n
        cv = 2 * h * (v0 @ n)
        mu_eff = 1.0 if mu * abs(cv) > 0.5 * (vt @ vt) else mu
        assert (mu_eff == 1.0) == static
        dv_t = -0.5 * friction_scale(w, vt) * vt
        np.testing.assert_allclose(v, v0 + (dv_n + mu_eff * dv_t) / 2, rtol=1e-10, atol=1e-14)


def test_vt_st_separating_vertex_untouched(orc):
    rec, v = vt_st_case(orc, np.array([0.0, 0.005, 0.0]), np.array([0.3, 0.2, 0.1]))
    assert rec["num"][0] == 1
    np.testing.assert_array_equal(v, [0.3, 0.2, 0.1])      # Cv >= 0: no delta, nc = 0


def test_vt_st_moving_static_triangle(orc):
    """The static velocities enter through the barycentric point velocity:
    Cv = g0.(v0 - vI), friction toward vI (vI = sum b_k vs_k)."""
    h, w, mu = 0.004, 1.5, 0.5
    vs = np.array([[0.2, 0.0, 0.1], [-0.1, 0.05, 0.0], [0.3, -0.02, 0.2]])
    b = np.array([0.25, 0.25, 0.5])
    vi = b @ vs
    v0 = np.array([0.1, -0.3, -0.2])
    rec, v = vt_st_case(orc, np.array([0.0, h, 0.0]), v0, w=w, mu=mu, vs=vs)
    n = np.array([0.0, 1.0, 0.0])
    rel = v0 - vi
    dv_n = -(rel @ n) * n
    vt = v0 + dv_n
    gt = vi - vt
    cv = 2 * h * (rel @ n)
    mu_eff = 1.0 if mu * abs(cv) > 0.5 * (gt @ gt) else mu
    dv_t = 0.5 * friction_scale(w, gt) * gt
    np.testing.assert_allclose(v, v0 + (dv_n + mu_eff * dv_t) / 2, rtol=1e-10, atol=1e-14)


def edge_point(a, b, away, delta=0.004, eps=0.003):
    """A point delta outside the midpoint of edge (a, b) in the plane (direction
    away) and eps above it, and that midpoint."""
    mid = 0.5 * (a + b)
    return mid + delta * unit(away) + np.array([0.0, eps, 0.0]), mid


def test_vt_st_type5_friction_sign(orc):
    """Edge (t2, t0): the friction delta points along vT (g0Tan = vTan0 - p,
    collision_constraints_v.py:109), so the tangential speed grows."""
    w, mu = 2.0, 100.0
    p, mid = edge_point(TRI[2], TRI[0], np.array([-2.0, 0.0, 1.0]))
    v0 = np.array([0.3, -0.2, 0.1])
    rec, v = vt_st_case(orc, p, v0, w=w, mu=mu)
    assert rec["dtype"][0, 0] == 5
    n = unit(p - mid)
    np.testing.assert_allclose(rec["g"][0, 0, 0], 2 * (p - mid), rtol=1e-10)
    dv_n = -(v0 @ n) * n
    vt = v0 + dv_n
    dv_t = +0.5 * friction_scale(w, vt) * vt                 # sign as the reference
    np.testing.assert_allclose(v, v0 + (dv_n + dv_t) / 2, rtol=1e-10, atol=1e-14)
    vt_new = v - (v @ n) * n
    assert np.linalg.norm(vt_new) > np.linalg.norm(vt)


def test_vt_st_type3_uses_static_inverse_mass(orc):
    """Edge (t0, t1): the normal delta uses the static mesh's m_inv at the
    dynamic vertex's index (collision_constraints_v.py:64); 0 past the static
    vertex count; friction keeps the dynamic inverse mass."""
    w, mu = 2.0, 0.8
    p, mid = edge_point(TRI[0], TRI[1], np.array([0.0, 0.0, -1.0]))
    v0 = np.array([0.1, -0.2, 0.3])
    n = unit(p - mid)
    for m_st, idx, ratio in ((np.array([5 * w, 1.0, 1.0]), 0, 5.0), (None, 0, 0.0),
                             (np.array([5 * w, 1.0, 1.0]), 3, 0.0)):
        rec, v = vt_st_case(orc, p, v0, w=w, mu=mu, m_inv_st=m_st, dyn_index=idx)
        assert rec["dtype"][idx, 0] == 3
        cv = rec["g"][idx, 0, 0] @ v0
        dv_n = -ratio * (v0 @ n) * n
        vt = v0 - (v0 @ n) * n                               # vTan0 with the dynamic mass
        mu_eff = 1.0 if mu * abs(cv) > 0.5 * (vt @ vt) else mu
        dv_t = -0.5 * friction_scale(w, vt) * vt
        np.testing.assert_allclose(v, v0 + (dv_n + mu_eff * dv_t) / 2, rtol=1e-10, atol=1e-14)


def tv_st_case(orc, p_static, vel, *, m_inv=(1.0, 1.0, 1.0), mu=0.8, fixed=(1.0, 1.0, 1.0)):
    """A static vertex against the dynamic triangle TRI (vertices 0, 1, 2)."""
    faces = np.array([[0, 1, 2]], np.int32)
    mesh = dict(faces=faces)
    y = TRI.copy()
    xs = np.array([p_static])
    cand = (np.array([[0]], np.int32), np.array([1], np.int32))
    fx = np.array(fixed, np.float64)
    mi = np.array(m_inv, np.float64)
    out = orc.collision_pass(mesh, np.zeros((1, 3), np.int32), y, xs, fx, mi, DHAT, 1e8, 4, None, None, cand,
                             record=True)
    v = np.array(vel, np.float64)
    v_in = v.copy()
    orc.velocity_pass(mesh, np.zeros((1, 3), np.int32), v, None, fx, mi, None, mu, out["pairs"], 4)
    return out["pairs"][2], v_in, v


def test_tv_st_type1_has_no_velocity_response(orc):
    """Static point at t1: g1 = g3 = 0 and Cv = g0.vs + g1.v1 + g3.v2 + g3.v3 = 0
    (g3 in place of g2, collision_constraints_v.py:144), so an approaching t1
    is left alone."""
    d, e = 0.003, 0.002
    p = TRI[1] + np.array([d, e, -d])
    rec, v_in, v = tv_st_case(orc, p, [[0, 0, 0], [0.5, 0.1, -0.5], [0, 0, 0]])
    assert rec["num"][0] == 1 and rec["dtype"][0, 0] == 1
    g = rec["g"][0, 0]
    assert g[2] @ v_in[1] < 0                          # t1 approaches the static point
    np.testing.assert_array_equal(v, v_in)


def test_tv_st_type2_uses_g3_for_t1(orc):
    """Static point at t2: Cv = g3.(v_t1 + v_t2).  With t2 at rest and t1 moving
    toward the point, t2 (and only t2) receives the normal and friction deltas."""
    d, e, w, mu = 0.004, 0.003, 3.0, 0.8
    p = TRI[2] + np.array([0.0, e, d])
    vel = [[0, 0, 0], [0.0, 0.3, 0.4], [0, 0, 0]]
    rec, v_in, v = tv_st_case(orc, p, vel, m_inv=(1.0, 1.0, w), mu=mu)
    assert rec["dtype"][0, 0] == 2
    g3 = rec["g"][0, 0, 3]
    np.testing.assert_allclose(g3, -2 * (p - TRI[2]), rtol=1e-12)
    n = unit(g3)
    dv_n = -(v_in[1] @ n) * n
    cv = g3 @ v_in[1]
    dd = dv_n @ dv_n
    mu_eff = 1.0 if mu * abs(cv) > 0.5 * dd else mu
    dv_t = -0.5 * friction_scale(w, dv_n) * dv_n
    np.testing.assert_allclose(v[2], (dv_n + mu_eff * dv_t) / 2, rtol=1e-10, atol=1e-14)
    np.testing.assert_array_equal(v[:2], v_in[:2])


def test_tv_st_face_conserves_triangle_momentum_direction(orc):
    """Static point above the face of a translating triangle: after the pass the
    triangle's barycentric point no longer approaches the point (normal part)
    and every vertex's delta is parallel to n plus the common friction term."""
    h = 0.004
    p = np.array([0.0, h, 0.0])
    u = np.array([0.1, 0.25, -0.05])                    # whole triangle moves up, toward p
    rec, v_in, v = tv_st_case(orc, p, [u, u, u], mu=100.0)
    assert rec["dtype"][0, 0] == 6
    b = np.array([0.25, 0.25, 0.5])
    dv = v - v_in
    # normal part of each vertex's delta is proportional to its barycentric weight
    np.testing.assert_allclose(dv[:, 1] / b, dv[0, 1] / b[0], rtol=1e-10)
    assert dv[0, 1] < 0


def vt_dy_case(orc, p, t, vel, m_inv, mu):
    """Dynamic vertex 0 against dynamic face (1, 2, 3) = triangle t."""
    faces = np.array([[1, 2, 3]], np.int32)
    mesh = dict(faces=faces)
    y = np.vstack([p, t])
    cand = np.full((4, 2), -1, np.int32)
    cnt = np.zeros(4, np.int32)
    cand[0, 0], cnt[0] = 0, 1
    fixed = np.ones(4)
    mi = np.array(m_inv, np.float64)
    out = orc.collision_pass(mesh, None, y, None, fixed, mi, DHAT, 1e8, 2, None, (cand, cnt), None, record=True)
    v = np.array(vel, np.float64)
    v_in = v.copy()
    orc.velocity_pass(mesh, None, v, None, fixed, mi, None, mu, out["pairs"], 2)
    return out["pairs"][1], v_in, v


def test_vt_dy_point_point_conserves_momentum(orc):
    """Equal masses, point-point self contact: deltas are equal and opposite."""
    d, e = 0.003, 0.002
    p = TRI[0] + np.array([-d, e, -d])
    vel = [[0.4, -0.1, 0.3], [-0.2, 0.1, 0.0], [0, 0, 0], [0, 0, 0]]
    for mu in (0.8, 100.0):
        rec, v_in, v = vt_dy_case(orc, p, TRI, vel, [2.0, 2.0, 1.0, 1.0], mu)
        assert rec["dtype"][0, 0] == 0
        np.testing.assert_allclose(v[0] + v[1], v_in[0] + v_in[1], rtol=1e-10, atol=1e-14)
        np.testing.assert_array_equal(v[2:], v_in[2:])
        assert np.linalg.norm(v[0] - v[1]) < np.linalg.norm(v_in[0] - v_in[1])


def test_vt_dy_type1_friction_uses_t0_inverse_mass(orc):
    """Point-point with t1 (face vertex 2): t1's friction delta is scaled by the
    inverse mass of t0 (face vertex 1), collision_constraints_v.py:388."""
    d, e, mu = 0.003, 0.002, 100.0
    p = TRI[1] + np.array([d, e, -d])
    m = np.array([2.0, 7.0, 3.0, 1.0])
    vel = np.array([[-0.5, -0.1, 0.5], [0, 0, 0], [0.1, 0.05, 0.2], [0, 0, 0]])
    rec, v_in, v = vt_dy_case(orc, p, TRI, vel, m, mu)
    assert rec["dtype"][0, 0] == 1
    g0 = rec["g"][0, 0, 0]
    np.testing.assert_allclose(g0, 2 * (p - TRI[1]), rtol=1e-12)
    schur = (m[0] + m[2]) * (g0 @ g0) + EPS
    assert np.isclose(rec["schur"][0, 0], schur, rtol=1e-12, atol=0)
    cv = g0 @ (v_in[0] - v_in[2])
    ld = -cv / schur
    dv0_n, dv2_n = m[0] * ld * g0, -m[2] * ld * g0
    r = (v_in[0] + dv0_n) - (v_in[2] + dv2_n)
    lt = (r @ r) / ((m[0] + m[2]) * (r @ r) + EPS)
    mu_eff = 1.0 if mu * abs(cv) > r @ r else mu
    np.testing.assert_allclose(v[0], v_in[0] + (dv0_n - mu_eff * m[0] * lt * r) / 2, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(v[2], v_in[2] + (dv2_n + mu_eff * m[1] * lt * r) / 2, rtol=1e-10, atol=1e-14)


def test_vt_dy_face_moves_point_and_face_apart(orc):
    """Point above a face, approaching: afterwards the point's velocity relative
    to the face's barycentric velocity has no approaching normal component left
    up to the 1e-4 regulariser, and unit-mass momentum is conserved."""
    h = 0.004
    p = np.array([0.0, h, 0.0])
    vel = [[0.1, -0.3, 0.05], [0, 0.1, 0], [0, 0.1, 0], [0, 0.1, 0]]
    rec, v_in, v = vt_dy_case(orc, p, TRI, vel, [1.0, 1.0, 1.0, 1.0], 100.0)
    assert rec["dtype"][0, 0] == 6
    np.testing.assert_allclose(v.sum(axis=0), v_in.sum(axis=0), rtol=1e-10, atol=1e-14)


def test_update_dv_pinned_and_massless(orc):
    """update_dv (XPBD.py:397-404): a pinned vertex (f = 0) and one with m_inv = 0
    keep their velocity; without pairs nothing changes."""
    h = 0.005
    faces_st = np.array([[0, 1, 2]], np.int32)
    mesh = dict(faces=np.zeros((1, 3), np.int32))
    for fixed, m in ((0.0, 2.0), (1.0, 0.0)):
        y = np.array([[0.0, h, 0.0]])
        cand = (np.array([[0]], np.int32), np.array([1], np.int32))
        fx, mi = np.array([fixed]), np.array([m])
        out = orc.collision_pass(mesh, faces_st, y, TRI.copy(), fx, mi, DHAT, 1e8, 4, cand, None, None, record=True)
        v = np.array([[0.3, -0.2, 0.1]])
        r = orc.velocity_pass(mesh, faces_st, v, None, fx, mi, None, 0.8, out["pairs"], 4)
        np.testing.assert_array_equal(v, [[0.3, -0.2, 0.1]])
        if m > 0:
            assert r["nc"][0] == 2
    v = np.array([[0.3, -0.2, 0.1]])
    orc.velocity_pass(mesh, faces_st, v, None, np.ones(1), np.ones(1), None, 0.8, None, 4)
    np.testing.assert_array_equal(v, [[0.3, -0.2, 0.1]])


def test_records_capped_and_out_of_range_recorded_empty(orc):
    """tv_st records every candidate (zero gradients and schur when d >= dhat)
    until cap records exist; vt_st records contacts only."""
    faces = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.int32)
    mesh = dict(faces=faces)
    y = np.vstack([TRI, TRI + [0, 1, 0], TRI + [0, 2, 0]])
    xs = np.array([[0.0, 0.004, 0.0]])
    cand = (np.array([[1, 0, 2]], np.int32), np.array([3], np.int32))
    n = 9
    out = orc.collision_pass(mesh, np.zeros((1, 3), np.int32), y, xs, np.ones(n), np.ones(n), DHAT, 1e8, 2,
                             None, None, cand, record=True)
    rec = out["pairs"][2]
    assert rec["num"][0] == 2
    assert list(rec["fid"][0]) == [1, 0]
    assert rec["schur"][0, 0] == 0 and not rec["g"][0, 0].any()       # face 1 is 1 m away
    assert rec["schur"][0, 1] > 0 and rec["dtype"][0, 1] == 6


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_velocity_frame_loop_rests_on_obstacle(orc, slx, dtype):
    """OracleCloth with the velocity pass: a sheet dropped onto a static
    triangle stays above it (the velocity pass removes the approaching normal
    velocity every substep) and stays finite."""
    g = slx.scene.grid_cloth(6, size=0.4, height=0.01)
    m = slx.mesh.Mesh.from_trimesh(g, device=None)
    st = dict(x=(TRI * 2).astype(np.float32), faces=np.array([[0, 1, 2]], np.int32))
    sim = orc.OracleCloth(m.host_dict(), st, dhat=1e-4, stiffness_stretch=1e4, stiffness_bending=1e2,
                          g=[0, -9.8, 0], dt=1 / 60, dtype=dtype, padding=0.02, collision=True, velocity=True)
    for _ in range(3):
        sim.forward(5)
    assert np.isfinite(sim.x).all() and np.isfinite(sim.v).all()
    assert sim.x[:, 1].min() > -0.01
