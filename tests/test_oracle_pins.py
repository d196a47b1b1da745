"""Pins of the oracle pieces that the closed forms of test_oracle_jacobi.py leave
open, each chosen so that a plausible mistake in the restatement fails it:

* bending constraints act on x (not y) with the bending slot's stiffness
  (XPBD.py:164-184, the branch `bi = i - max_num_edges_dy` reads verts.x);
* the stiffness names are swapped when the solver calls the spring kernel:
  stretch constraints get stiffness_bending * dt^2 and bending constraints
  stiffness_stretch * dt^2 (XPBD.py:354-355);
* compute_velocity's damping factor (1 - damping) (XPBD.py:340-343), with
  damping != 0;
* the graph-coloured Gauss-Seidel pass: colours in order, each edge projected
  on the current y with nc = 1 (the sweep the colouring of
  graph_coloring.py:131-201 is built for), after one Jacobi-averaged bending
  pass on x.

Expected values are written out from the XPBD formulas by hand in float64,
not by calling the oracle a second way."""
import numpy as np
import pytest


def line_mesh(edges, l0, bend=(), bl0=()):
    return dict(edges=np.array(edges, np.int32).reshape(-1, 2), l0=np.array(l0, np.float64),
                bend=np.array(bend, np.int32).reshape(-1, 2), bl0=np.array(bl0, np.float64))


def xpbd_ld(k, C, w0, w1):
    return k * C / (k * (w0 + w1) + 1.0)


# ------------------------------------------------------------- bending on x

def test_bending_acts_on_x_with_the_bending_stiffness(orc):
    """One bending pair (0, 1), rest length 1: x stretched to 1.3 apart along
    +x while y sits exactly at the rest length.  The correction comes from x
    (C = 0.3) with k_b; a restatement reading y would see C = 0 and not move,
    one using k_s would move by a different amount."""
    k_s, k_b, w0, w1 = 999.0, 0.5, 2.0, 0.5
    mesh = line_mesh([], [], bend=[[0, 1]], bl0=[1.0])
    x = np.array([[0.0, 0, 0], [1.3, 0, 0]])
    y = np.array([[5.0, 1, 0], [6.0, 1, 0]])          # at rest length, elsewhere
    y0 = y.copy()
    orc.spring_pass(mesh, x, y, np.ones(2), np.array([w0, w1]), k_s, k_b)
    ld = xpbd_ld(k_b, 0.3, w0, w1)
    # n = (x0 - x1)/|x0 - x1| = -e_x: vertex 0 moves by +w0 ld, vertex 1 by -w1 ld
    np.testing.assert_allclose(y[0], y0[0] + [w0 * ld, 0, 0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(y[1], y0[1] - [w1 * ld, 0, 0], rtol=0, atol=1e-14)
    assert abs(ld - xpbd_ld(k_s, 0.3, w0, w1)) > 1e-3      # the slots are distinguishable


def test_stretch_on_y_and_bending_on_x_share_the_jacobi_average(orc):
    """Stretch (0, 1) on y and bending (1, 2) on x meet at vertex 1, whose
    position moves by the mean of the two deltas (nc = 2); the ends move by
    their single delta."""
    k_s, k_b = 7.0, 3.0
    w = np.array([1.0, 2.0, 4.0])
    mesh = line_mesh([[0, 1]], [1.0], bend=[[1, 2]], bl0=[1.0])
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.5, 0, 0]])          # bending C = 0.5 on x
    y = np.array([[0.0, 0, 0], [1.2, 0, 0], [2.0, 0, 0]])          # stretch C = 0.2 on y
    y0 = y.copy()
    orc.spring_pass(mesh, x, y, np.ones(3), w, k_s, k_b)
    ls = xpbd_ld(k_s, 0.2, w[0], w[1])       # n = -e_x for both constraints
    lb = xpbd_ld(k_b, 0.5, w[1], w[2])
    d0 = +w[0] * ls
    d1 = (-w[1] * ls + w[1] * lb) / 2.0
    d2 = -w[2] * lb
    np.testing.assert_allclose(y[:, 0] - y0[:, 0], [d0, d1, d2], rtol=0, atol=1e-14)
    np.testing.assert_array_equal(y[:, 1:], 0.0)


# -------------------------------------------------- stiffness name swap

def two_vertex_cloth(orc, stiffness_stretch, stiffness_bending, damping=0.0, dt=1 / 60, stretch=1.3,
                     w=(2.0, 0.5)):
    x = np.array([[0.0, 0, 0], [stretch, 0, 0]])
    d = dict(x=x, x0=np.array([[0.0, 0, 0], [1.0, 0, 0]]), faces=np.zeros((0, 3), np.int32),
             edges=np.array([[0, 1]], np.int32), l0=np.array([1.0]), bend=np.zeros((0, 2), np.int32),
             bl0=np.zeros(0), fixed=np.ones(2), m_inv=np.array(w))
    return orc.OracleCloth(d, None, dhat=1e-4, stiffness_stretch=stiffness_stretch,
                           stiffness_bending=stiffness_bending, g=(0, 0, 0), dt=dt, dtype=np.float64,
                           damping=damping)


def test_stretch_constraint_uses_the_bending_stiffness(orc):
    """OracleCloth, one stretch edge stretched by 0.3, no gravity, no damping,
    one substep of one iteration: x_new = y = x + w ld n with
    ld = k C / (k S + 1) and k = stiffness_bending * dt^2 (XPBD.py:354-355)."""
    dt, ks, kb, w0, w1 = 1 / 60, 1e6, 50.0, 2.0, 0.5
    sim = two_vertex_cloth(orc, ks, kb, dt=dt)
    sim.forward(1, 1)
    ld = xpbd_ld(kb * dt * dt, 0.3, w0, w1)
    np.testing.assert_allclose(sim.x[:, 0], [w0 * ld, 1.3 - w1 * ld], rtol=0, atol=1e-13)
    wrong = xpbd_ld(ks * dt * dt, 0.3, w0, w1)
    assert abs(sim.x[0, 0] - w0 * wrong) > 1e-3
    # the velocity the substep leaves: (y - x) / dt with no damping
    np.testing.assert_allclose(sim.v[:, 0], [w0 * ld / dt, -w1 * ld / dt], rtol=1e-12)


def test_bending_constraint_uses_the_stretch_stiffness(orc):
    """The other half of the swap: a bending pair stretched on x is corrected
    with k = stiffness_stretch * dt^2."""
    dt, ks, kb = 1 / 60, 30.0, 1e5
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.4, 0, 0]])
    d = dict(x=x, x0=np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]]), faces=np.zeros((0, 3), np.int32),
             edges=np.zeros((0, 2), np.int32), l0=np.zeros(0), bend=np.array([[0, 2]], np.int32),
             bl0=np.array([2.0]), fixed=np.ones(3), m_inv=np.ones(3))
    sim = orc.OracleCloth(d, None, dhat=1e-4, stiffness_stretch=ks, stiffness_bending=kb, g=(0, 0, 0), dt=dt,
                          dtype=np.float64, damping=0.0)
    sim.forward(1, 1)
    ld = xpbd_ld(ks * dt * dt, 0.4, 1.0, 1.0)
    np.testing.assert_allclose(sim.x[:, 0], [ld, 1.0, 2.4 - ld], rtol=0, atol=1e-13)


# ------------------------------------------------------------- damping

def test_damped_free_fall_closed_form(slx, orc):
    """No constraint active (rest pose, no pins, uniform translation), damping
    d: v_k = a (v_{k-1} + g h) with a = 1 - d, x_k = x_{k-1} + h v_k, so
    x_n - x_0 = g h^2 a/(1-a) (n - a (1 - a^n)/(1 - a)).  Without the factor
    (or with it applied to x instead of v) the fall differs."""
    gc = slx.scene.grid_cloth(6, jitter=0.0)
    m = slx.Mesh(x=gc.x, faces=gc.faces, device=None)
    d = m.host_dict()
    d["fixed"] = np.ones(m.num_verts, np.float32)
    damping, n = 0.1, 12
    sim = orc.OracleCloth(d, None, dhat=1e-4, stiffness_stretch=1e6, stiffness_bending=1e3,
                          g=(0, -9.8, 0), dt=1 / 60, dtype=np.float64, damping=damping)
    sim.forward(n, 1)
    h = (1 / 60) / n
    a = 1.0 - float(np.float32(damping))         # damping is a float32 kernel argument
    g32 = -float(np.float32(9.8))
    expect = g32 * h * h * a / (1 - a) * (n - a * (1 - a ** n) / (1 - a))
    np.testing.assert_allclose(sim.x[:, 1] - d["x"][:, 1].astype(np.float64), expect, rtol=1e-9)
    vexp = g32 * h * a * (1 - a ** n) / (1 - a)
    np.testing.assert_allclose(sim.v[:, 1], vexp, rtol=1e-9)
    undamped = g32 * h * h * n * (n + 1) / 2
    assert abs(expect - undamped) > 1e-3 * abs(undamped)


def test_damping_scales_the_spring_velocity(orc):
    """compute_velocity after a spring correction: v = (1 - d)(y - x)/dt and
    x moves by dt v, i.e. by (1 - d) of the correction."""
    dt, kb, damping = 1 / 60, 50.0, 0.25
    sim = two_vertex_cloth(orc, 1e6, kb, damping=damping, dt=dt)
    sim.forward(1, 1)
    ld = xpbd_ld(kb * dt * dt, 0.3, 2.0, 0.5)
    a = 1.0 - float(np.float32(damping))
    np.testing.assert_allclose(sim.x[:, 0], [a * 2.0 * ld, 1.3 - a * 0.5 * ld], rtol=0, atol=1e-13)


# ------------------------------------------------------------- gs_pass

def test_gs_single_edge_equals_the_two_particle_form(orc):
    w0, w1, k = 2.0, 0.5, 50.0
    mesh = line_mesh([[0, 1]], [1.0])
    x = np.array([[0.0, 0, 0], [1.3, 0, 0]])
    y = x.copy()
    orc.gs_pass(mesh, x, y, np.ones(2), np.array([w0, w1]), k, 0.0, np.array([0]), np.array([0, 1]))
    ld = xpbd_ld(k, 0.3, w0, w1)
    np.testing.assert_allclose(y[:, 0], [w0 * ld, 1.3 - w1 * ld], rtol=0, atol=1e-14)


@pytest.mark.parametrize("swap_colours", [False, True])
def test_gs_chain_in_two_colours(orc, swap_colours):
    """Chain 0-1-2 on the x axis, unit masses, rest lengths 1, y = (0, 1.3, 2.7).
    Colour order (0, 1): edge (0,1) first, C0 = 0.3, l0 = k C0/(2k+1): y0 += l0,
    y1 -= l0; then edge (1,2) sees the moved y1: C1 = 0.4 + l0, l1 = k C1/(2k+1):
    y1 += l1, y2 -= l1.  The opposite colour order gives a different answer, and
    so would Jacobi averaging."""
    k = 5.0
    mesh = line_mesh([[0, 1], [1, 2]], [1.0, 1.0])
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    y = np.array([[0.0, 0, 0], [1.3, 0, 0], [2.7, 0, 0]])
    order, off = (np.array([1, 0]), np.array([0, 1, 2])) if swap_colours else (np.array([0, 1]), np.array([0, 1, 2]))
    orc.gs_pass(mesh, x, y, np.ones(3), np.ones(3), k, 0.0, order, off)
    if not swap_colours:
        l0 = k * 0.3 / (2 * k + 1)
        l1 = k * (0.4 + l0) / (2 * k + 1)
        expect = [l0, 1.3 - l0 + l1, 2.7 - l1]
    else:
        l1 = k * 0.4 / (2 * k + 1)
        l0 = k * (0.3 + l1) / (2 * k + 1)
        expect = [l0, 1.3 + l1 - l0, 2.7 - l1]
    np.testing.assert_allclose(y[:, 0], expect, rtol=0, atol=1e-14)
    jac = np.array([0.0, 1.3, 2.7])
    ja, jb = k * 0.3 / (2 * k + 1), k * 0.4 / (2 * k + 1)
    jac += [ja, (-ja + jb) / 2, -jb]
    assert np.abs(y[:, 0] - jac).max() > 1e-3


def test_gs_bending_pass_on_x_precedes_the_colours(orc):
    """Bending pair (0, 2) on x (C = 0.4, k_b) is applied first as a Jacobi pass
    (nc = 1 for both ends), then the single stretch edge (0, 1) of colour 0 sees
    the bent y0."""
    k_s, k_b = 4.0, 9.0
    mesh = line_mesh([[0, 1]], [1.0], bend=[[0, 2]], bl0=[2.0])
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.4, 0, 0]])
    y = np.array([[0.0, 0, 0], [1.3, 0, 0], [2.4, 0, 0]])
    orc.gs_pass(mesh, x, y, np.ones(3), np.ones(3), k_s, k_b, np.array([0]), np.array([0, 1]))
    lb = k_b * 0.4 / (2 * k_b + 1)
    y0, y2 = lb, 2.4 - lb
    C = 1.3 - y0 - 1.0
    ls = k_s * C / (2 * k_s + 1)
    np.testing.assert_allclose(y[:, 0], [y0 + ls, 1.3 - ls, y2], rtol=0, atol=1e-14)
