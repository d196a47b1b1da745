"""Mesh topology for the XPBD cloth step (host side, numpy).

Everything here runs once per mesh, before any substep, and produces the
arrays the HIP kernels read:

* canonical edge list (stretch constraints),
* bending pairs in the reference's emission order
  (framework/meshio/meshTaichiWrapper.py:152-225),
* rest lengths and the lumped inverse mass
  (framework/meshio/meshTaichiWrapper.py:141-145, 243-256),
* the per-vertex constraint adjacency in SELL-64 layout, which lets one
  thread per vertex gather its Jacobi delta in the same order in which a serial
  scatter over constraints (framework/physics/XPBD.py:140-184) would have
  accumulated it.
"""
from __future__ import annotations

import numpy as np

SLICE = 64          # rows per SELL slice = lanes per wavefront
BEND_BIT = np.int64(1) << 31


def edges_from_faces(faces: np.ndarray) -> np.ndarray:
    """Unique undirected edges of a triangle mesh, (lo, hi) per row, sorted
    lexicographically.  This is the canonical stretch-constraint order of this
    package (meshtaichi_patcher's own edge numbering is not reproducible)."""
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
    e.sort(axis=1)
    e = np.unique(e, axis=0)
    return e.astype(np.int32)


def face_edges(faces: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Edge ids of each face's sides (v0v1, v1v2, v2v0) in `edges` (lexicographic
    (lo, hi) rows, as edges_from_faces gives): the reference's
    face_edge_indices (framework/meshio/meshTaichiWrapper.py:115-119), whose
    order within a face follows meshtaichi's relation order; this package fixes
    it to the sides in vertex order."""
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if f.shape[0] == 0:
        return np.zeros((0, 3), np.int32)
    n = int(max(f.max(), e.max() if e.size else 0)) + 1
    key = e[:, 0] * n + e[:, 1]
    if np.any(key[1:] <= key[:-1]):
        raise ValueError("edges must be unique (lo, hi) rows in lexicographic order")
    sides = np.stack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=1)     # [F, 3, 2]
    lo, hi = sides.min(axis=2), sides.max(axis=2)
    q = lo * n + hi
    at = np.searchsorted(key, q)
    at = np.minimum(at, len(key) - 1)
    if not np.all(key[at] == q):
        raise ValueError("a face side is missing from the edge list")
    return at.astype(np.int32)


def edge_order(edges: np.ndarray, vert_order: np.ndarray) -> np.ndarray:
    """A spatially coherent processing order of the edges: by the position of
    their first vertex in vert_order (a Morton order of the vertices), stable."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    rank = np.empty(len(vert_order), np.int64)
    rank[np.asarray(vert_order, np.int64)] = np.arange(len(vert_order))
    return np.argsort(rank[e[:, 0]], kind="stable").astype(np.int32)


def own_faces(nv: int, faces: np.ndarray, width: int = 8) -> np.ndarray:
    """int32 [nv, width]: the faces incident to each vertex, in increasing face
    id, -1 padded.  A vertex on more than `width` faces gets a row of -1 (its
    candidate lists then keep its own faces; the narrow phase skips them)."""
    fid = np.asarray(faces, np.int64).reshape(-1, 3)
    out = np.full((nv, width), -1, np.int32)
    if len(fid) == 0 or nv == 0:
        return out
    v = fid.reshape(-1)
    f = np.repeat(np.arange(len(fid), dtype=np.int64), 3)
    o = np.lexsort((f, v))
    v, f = v[o], f[o]
    first = np.ones(len(v), bool)              # a face listing a vertex twice counts once
    first[1:] = (v[1:] != v[:-1]) | (f[1:] != f[:-1])
    v, f = v[first], f[first]
    cnt = np.bincount(v, minlength=nv)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rank = np.arange(len(v)) - start[v]
    ok = cnt[v] <= width
    out[v[ok], rank[ok]] = f[ok]
    return out


def bending_pairs(faces: np.ndarray) -> np.ndarray:
    """Opposite-vertex pairs across interior edges.

    Restates meshTaichiWrapper.findTriNeighbors (:152-193) and getBendingPair
    (:195-225): every face contributes local edges (a,b),(a,c),(b,c) of its
    sorted vertices with ids 3f+k; rows are stably sorted by (lo, hi); equal
    consecutive rows are paired greedily; pairs are emitted at the smaller of
    the two local edge ids, in ascending order; the pair is (vertex of f not in
    the neighbour, vertex of the neighbour not in f)."""
    f = np.sort(np.asarray(faces, dtype=np.int64).reshape(-1, 3), axis=1)
    nf = f.shape[0]
    if nf == 0:
        return np.zeros((0, 2), np.int32)
    lo = np.stack([f[:, 0], f[:, 0], f[:, 1]], axis=1).reshape(-1)
    hi = np.stack([f[:, 1], f[:, 2], f[:, 2]], axis=1).reshape(-1)
    eid = np.arange(3 * nf, dtype=np.int64)
    order = np.lexsort((hi, lo))                # stable, primary lo, then hi
    slo, shi, seid = lo[order], hi[order], eid[order]
    n = slo.shape[0]
    # greedy walk over runs of equal keys pairs positions (0,1),(2,3),...
    new_run = np.ones(n, dtype=bool)
    new_run[1:] = (slo[1:] != slo[:-1]) | (shi[1:] != shi[:-1])
    run_start = np.maximum.accumulate(np.where(new_run, np.arange(n), 0))
    pos = np.arange(n) - run_start
    nxt_same = np.zeros(n, dtype=bool)
    nxt_same[:-1] = ~new_run[1:]
    first = (pos % 2 == 0) & nxt_same
    a = seid[first]
    b = seid[np.nonzero(first)[0] + 1]
    e_min = np.minimum(a, b)
    e_max = np.maximum(a, b)
    k = np.argsort(e_min, kind="stable")
    e_min, e_max = e_min[k], e_max[k]
    fa = f[e_min // 3]
    fb = f[e_max // 3]
    pa = _not_in(fa, fb)
    pb = _not_in(fb, fa)
    return np.stack([pa, pb], axis=1).astype(np.int32)


def _not_in(v: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Per row, the single entry of v (sorted triple) absent from w."""
    m = (v[:, :, None] == w[:, None, :]).any(axis=2)
    if not np.all((~m).sum(axis=1) == 1):
        raise ValueError("bending pair across a degenerate face pair")
    return v[~m]


def norm_f32(d: np.ndarray) -> np.ndarray:
    """|d| in float32 with the op order sqrt((dx*dx + dy*dy) + dz*dz)."""
    d = d.astype(np.float32)
    s = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    s = s + d[:, 2] * d[:, 2]
    return np.sqrt(s).astype(np.float32)


def rest_lengths(x: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    x = np.asarray(x, np.float32)
    return norm_f32(x[pairs[:, 0]] - x[pairs[:, 1]])


def lumped_inverse_mass(nv: int, edges: np.ndarray, l0: np.ndarray) -> np.ndarray:
    """m_inv[v] = 1 / sum_{edges e containing v} 0.5*l0[e], accumulated in
    float32 in edge order, endpoint 0 before endpoint 1
    (meshTaichiWrapper.py:243-256).  Isolated vertices get +inf like the
    reference's 1/0."""
    m = np.zeros(nv, np.float32)
    half = (np.float32(0.5) * l0.astype(np.float32)).astype(np.float32)
    idx = np.asarray(edges, np.int64).reshape(-1)            # e0v0,e0v1,e1v0,...
    vals = np.repeat(half, 2)
    np.add.at(m, idx, vals)
    with np.errstate(divide="ignore"):
        return (np.float32(1.0) / m).astype(np.float32)


class Adjacency:
    """Per-vertex constraint adjacency in SELL-64 layout.

    Row v lists, in ascending global constraint index c (stretch edges
    c in [0,E) first, then bending pairs c in [E,E+B)), the other endpoint u
    and the rest length.  Entry word 0 is u, with bit 31 set for a bending
    constraint; word 1 holds the float32 bits of the rest length.  Slice s
    covers rows [64 s, 64 s + 64); entry k of row v sits at
    slice_off[s] + k*64 + (v - 64 s).  Padding entries have word 0 = -1.
    """

    def __init__(self, nv: int, edges: np.ndarray, l0: np.ndarray,
                 bend: np.ndarray, bl0: np.ndarray):
        ne, nb = len(edges), len(bend)
        a = np.concatenate([edges[:, 0], bend[:, 0]]).astype(np.int64)
        b = np.concatenate([edges[:, 1], bend[:, 1]]).astype(np.int64)
        rl = np.concatenate([l0, bl0]).astype(np.float32)
        kind = np.concatenate([np.zeros(ne, np.int64), np.ones(nb, np.int64)])
        nc = ne + nb
        # two directed entries per constraint, in constraint order
        own = np.stack([a, b], axis=1).reshape(-1)
        oth = np.stack([b, a], axis=1).reshape(-1)
        rl2 = np.repeat(rl, 2)
        kind2 = np.repeat(kind, 2)
        order = np.argsort(own, kind="stable")
        own, oth, rl2, kind2 = own[order], oth[order], rl2[order], kind2[order]
        deg = np.bincount(own, minlength=nv).astype(np.int64)
        self.nv, self.nnz = nv, 2 * nc
        self.degree = deg.astype(np.int32)
        row_off = np.zeros(nv + 1, np.int64)
        np.cumsum(deg, out=row_off[1:])
        self.row_off = row_off
        self.other = oth.astype(np.int32)
        self.kind = kind2.astype(np.int32)
        self.rest = rl2
        # SELL-64 packing
        ns = (nv + SLICE - 1) // SLICE
        pad_deg = np.zeros(ns * SLICE, np.int64)
        pad_deg[:nv] = deg
        width = pad_deg.reshape(ns, SLICE).max(axis=1)
        slice_off = np.zeros(ns + 1, np.int64)
        np.cumsum(width * SLICE, out=slice_off[1:])
        total = int(slice_off[-1])
        ent = np.empty((total, 2), np.int32)
        ent[:, 0] = -1
        ent[:, 1] = 0
        k_in_row = np.arange(2 * nc, dtype=np.int64) - np.repeat(row_off[:-1], deg)
        rows = own
        pos = slice_off[rows // SLICE] + k_in_row * SLICE + (rows % SLICE)
        w0 = oth | (kind2 * BEND_BIT)
        ent[pos, 0] = (w0 & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
        ent[pos, 1] = rl2.view(np.int32)
        self.slice_off = slice_off.astype(np.int32)   # [ns+1], in entries
        self.slice_width = width.astype(np.int32)     # [ns]
        self.entries = ent                            # [total, 2] int32
        if total >= 2 ** 31:
            raise ValueError("adjacency exceeds int32 addressing")

    @property
    def num_slices(self) -> int:
        return len(self.slice_width)


def greedy_edge_coloring(nv: int, edges: np.ndarray, max_colors: int = 64):
    """Greedy colouring of constraints (edges conflict when they share a
    vertex), edges visited in index order, each taking the lowest colour not
    used by an already-coloured adjacent edge
    (framework/utilities/graph_coloring.py:131-149).  Returns (color[E],
    order[E], offsets[n_colors+1]) where order lists edge ids grouped by colour,
    ascending id within a colour (the stable counting sort of
    graph_coloring.py:172-201)."""
    from . import _abi
    return _abi.greedy_coloring(nv, edges, max_colors)


def _spread10(v):
    """Insert two zero bits between each of the low 10 bits (3-D Morton)."""
    v = v.astype(np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def morton_order(x):
    """Vertex permutation in 30-bit Morton order of the positions x (V,3) within
    their bounding box (ties by index).  Used only as the order in which
    traversal threads take their queries: lanes of a wavefront then query
    neighbouring points and share tree nodes; results do not depend on it."""
    x = np.asarray(x, np.float64)
    if len(x) == 0:
        return np.zeros(0, np.int32)
    lo, hi = x.min(axis=0), x.max(axis=0)
    ext = np.where(hi > lo, hi - lo, 1.0)
    q = np.clip(((x - lo) / ext * 1023.0).astype(np.int64), 0, 1023)
    code = (_spread10(q[:, 0]) << np.uint64(2)) | (_spread10(q[:, 1]) << np.uint64(1)) | _spread10(q[:, 2])
    return np.argsort(code, kind="stable").astype(np.int32)


def sewing_lists(pairs: np.ndarray, nv: int):
    """Per sewn vertex, its sewing pair entries in pair order (the order in which
    the reference's pair loop scatters into that vertex, XPBD.py:420-441):
    returns (vertices [n], offsets [n+1], entries [2 np] of 2 * pair + side)."""
    pr = np.ascontiguousarray(pairs, np.int64).reshape(-1, 2)
    if pr.size and (pr.min() < 0 or pr.max() >= nv):
        raise IndexError(f"sewing pair names a vertex outside [0, {nv})")
    ent_v = pr.reshape(-1)                       # entry 2p + side -> vertex
    ent = np.arange(ent_v.size, dtype=np.int64)
    order = np.lexsort((ent, ent_v))             # by vertex, then by pair order
    verts, starts = np.unique(ent_v[order], return_index=True)
    offs = np.concatenate([starts, [ent_v.size]]).astype(np.int32)
    return verts.astype(np.int32), offs, ent[order].astype(np.int32)
