"""Sewing constraints (XPBD.py:418-441) as a Jacobi pass, and the per-vertex
pair lists the device pass gathers from (CPU).

Closed forms: a free pair of inverse masses w1, w2 ends at separation
C / (k (w1 + w2) + 1) along the same line, each point moving in proportion to
its inverse mass (the pair's centre weighted by mass is kept); a pinned point
does not move; a vertex in two pairs moves by the mean of its two deltas; a
coincident pair stays put."""
import numpy as np


def one_pass(orc, y, pairs, fixed, m_inv, k):
    y = np.array(y, np.float64)
    orc.sewing_pass(y, np.array(pairs), np.array(fixed, np.float64), np.array(m_inv, np.float64), k)
    return y


def test_free_pair_closed_form(orc):
    y0 = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    w1, w2, k = 2.0, 0.5, 7.0
    y = one_pass(orc, y0, [[0, 1]], [1, 1], [w1, w2], k)
    C = np.linalg.norm(y0[1] - y0[0])
    np.testing.assert_allclose(np.linalg.norm(y[1] - y[0]), C / (k * (w1 + w2) + 1), rtol=1e-12)
    u = (y0[1] - y0[0]) / C
    np.testing.assert_allclose(np.cross(y[1] - y[0], u), 0, atol=1e-14)          # same line
    com0 = (y0[0] / w1 + y0[1] / w2) / (1 / w1 + 1 / w2)
    com = (y[0] / w1 + y[1] / w2) / (1 / w1 + 1 / w2)
    np.testing.assert_allclose(com, com0, atol=1e-14)


def test_pinned_point_and_shared_vertex(orc):
    y0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    k, w = 3.0, 1.0
    y = one_pass(orc, y0, [[0, 1]], [0, 1, 1], [w, w, w], k)
    np.testing.assert_array_equal(y[0], y0[0])
    np.testing.assert_allclose(y[1], [1.0 - k * w / (k * w + 1), 0, 0], rtol=1e-12)
    # vertex 0 (free now) in two pairs: it moves by the mean of the two pair deltas
    y = one_pass(orc, y0, [[0, 1], [2, 0]], [1, 1, 1], [w, w, w], k)
    ld1 = k * 1.0 / (k * 2 * w + 1)
    ld2 = k * 2.0 / (k * 2 * w + 1)
    d0 = (w * ld1 * np.array([1.0, 0, 0]) + w * ld2 * np.array([0, 1.0, 0])) / 2
    np.testing.assert_allclose(y[0], y0[0] + d0, rtol=1e-12, atol=1e-15)


def test_coincident_pair_is_finite(orc):
    y0 = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [1.0, 0.5, 0.5]])
    y = one_pass(orc, y0, [[0, 1], [1, 2]], [1, 1, 1], [1, 1, 1], 2.0)
    assert np.isfinite(y).all()
    np.testing.assert_array_equal(y[0], y0[0])                # only the zero-length pair touches it
    assert y[1, 0] > 0.5 and y[2, 0] < 1.0


def test_sewing_lists_follow_pair_order(slx, rng):
    nv = 40
    pairs = rng.integers(0, nv, (60, 2)).astype(np.int32)
    verts, offs, ent = slx.topology.sewing_lists(pairs, nv)
    expect = {}
    for p, (a, b) in enumerate(pairs):                        # the reference's loop, literally
        expect.setdefault(int(a), []).append(2 * p)
        expect.setdefault(int(b), []).append(2 * p + 1)
    assert list(verts) == sorted(expect)
    for i, v in enumerate(verts):
        assert list(ent[offs[i]:offs[i + 1]]) == expect[int(v)]
