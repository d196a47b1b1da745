"""Greedy stretch-edge colouring (graph_coloring.py:131-201) of the product
library against the oracle's restatement, plus the properties that pin the
greedy rule itself: validity (no two edges sharing a vertex share a colour),
minimality (every edge's colour is the lowest one not used by an earlier
adjacent edge) and the stable colour-major order."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def check_greedy(nv, edges, color, order, offsets):
    ne = len(edges)
    inc = [[] for _ in range(nv)]
    for e, (a, b) in enumerate(edges):
        inc[a].append(e)
        inc[b].append(e)
    for e, (a, b) in enumerate(edges):
        nb = set(inc[a]) | set(inc[b])
        nb.discard(e)
        assert all(color[o] != color[e] for o in nb)                   # valid
        used = {color[o] for o in nb if o < e}
        assert color[e] == min(set(range(len(used) + 1)) - used)       # greedy, lowest free
    assert sorted(order.tolist()) == list(range(ne))
    for c in range(len(offsets) - 1):
        cls = order[offsets[c]:offsets[c + 1]]
        assert np.all(color[cls] == c) and np.all(np.diff(cls) > 0)


@pytest.mark.parametrize("maker", ["grid16", "ico2", "plane8", "layers"])
def test_library_coloring_equals_oracle_and_is_greedy(slx, orc, maker):
    if maker == "grid16":
        t = slx.scene.grid_cloth(16)
    elif maker == "ico2":
        t = slx.scene.icosphere(2)
    elif maker == "plane8":
        t = slx.scene.read_obj(os.path.join(HERE, "golden", "plane_8.obj"))
    else:
        t = slx.scene.stacked_layers(9, 3, 0.02)
    e = slx.topology.edges_from_faces(t.faces)
    nv = len(t.x)
    c1, o1, f1 = slx.topology.greedy_edge_coloring(nv, e)
    c2, o2, f2 = orc.greedy_coloring(nv, e)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(f1, f2)
    check_greedy(nv, e, c1, o1, f1)


@pytest.mark.parametrize("n", [256, 1024])
def test_cloth_grids_need_at_most_16_colors(slx, n):
    t = slx.scene.grid_cloth(n)
    e = slx.topology.edges_from_faces(t.faces)
    _, _, off = slx.topology.greedy_edge_coloring(n * n, e)
    assert len(off) - 1 <= 16


def test_reference_precomputed_coloring_shape():
    """Weak fixture: the reference's cached colouring of plane_8.obj's 208 edges
    (patcher edge numbering, which is not reproducible here).  It must be a
    colour-major permutation with ascending ids per colour, and 208 = V + F - 1."""
    fx = json.load(open(os.path.join(HERE, "golden", "precomputed_coloring.json")))
    start = np.array(fx["color_start"])
    idx = np.array(fx["sorted_edges"])
    n_colors = int(np.argmax(start == start[-1]))
    assert start[n_colors] == 208 == 81 + 128 - 1
    assert sorted(idx.tolist()) == list(range(208))
    for c in range(n_colors):
        assert np.all(np.diff(idx[start[c]:start[c + 1]]) > 0)
    assert np.diff(start[:n_colors + 1]).tolist() == [33, 34, 24, 32, 25, 19, 16, 18, 7]


def test_coloring_rejects_too_few_colors(slx):
    t = slx.scene.grid_cloth(8)
    e = slx.topology.edges_from_faces(t.faces)
    with pytest.raises(slx.SlxError):
        slx.topology.greedy_edge_coloring(64, e, max_colors=3)
