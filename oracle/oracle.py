"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, as the checker.  The product path never imports
this module.

`OracleCloth` runs the reference's frame loop (framework/physics/XPBD.py:443-469)
serially on numpy arrays, in float32 or float64, with the same scalar
conversions as the reference (dt_sub, stiffness*dt^2 evaluated in double and
rounded to float32 when the arithmetic is float32).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
NCELLS = 32768


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _LIB = C.CDLL(path)
        _LIB.orc_morton3d.restype = C.c_int32
        _LIB.orc_expand_bits.restype = C.c_uint32
        _LIB.orc_expand_bits.argtypes = [C.c_uint32]
        _LIB.orc_morton3d.argtypes = [C.c_float] * 3
        _LIB.orc_flat_cell_id.restype = C.c_int32
        _LIB.orc_dist_f.restype = C.c_float
        _LIB.orc_dist_d.restype = C.c_double
        _LIB.orc_dist_ee_f.restype = C.c_float
        _LIB.orc_dist_ee_d.restype = C.c_double
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (the traversal, and the
    element loops of OracleCloth(parallel=True))."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def P(a):
    """Pointer of a contiguous numpy array (None stays NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _real(dtype, par=False):
    """(ctypes real, function suffix); par: the OpenMP instantiation (float32
    only, atomic scatter in no fixed order -- timing baseline, not a checker)."""
    if par:
        assert np.dtype(dtype) == np.float32, "the OpenMP instantiation is float32"
        return C.c_float, "p"
    return (C.c_float, "f") if np.dtype(dtype) == np.float32 else (C.c_double, "d")


# ------------------------------------------------------------------ primitives

def compute_y(x, v, fixed, dt, g, par=False):
    ct, s = _real(x.dtype, par)
    y = np.empty_like(x)
    gg = np.asarray(g, x.dtype)
    getattr(lib(), "orc_compute_y_" + s)(len(fixed), P(x), P(v), P(fixed), P(y), ct(dt), P(gg))
    return y


def spring_pass(mesh, x, y, fixed, m_inv, k_s, k_b, par=False):
    ct, s = _real(x.dtype, par)
    dt = x.dtype
    dx = np.zeros_like(x)
    nc = np.zeros(len(fixed), dt)
    l0 = mesh["l0"].astype(dt)
    bl0 = mesh["bl0"].astype(dt)
    getattr(lib(), "orc_spring_pass_" + s)(
        len(fixed), len(l0), len(bl0), P(mesh["edges"]), P(l0), P(mesh["bend"]), P(bl0),
        P(x), P(y), P(fixed), P(m_inv), ct(k_s), ct(k_b), P(dx), P(nc))
    return y


def implicit_pass(mesh, ya, y, fixed, m_inv, kappa):
    """One residual-solve pass (implicit-Euler stretch problem anchored at ya);
    updates y in place and returns the residual (largest step) of the input y."""
    ct, s = _real(y.dtype)
    gs = np.zeros_like(y)
    ks = np.zeros(len(fixed), y.dtype)
    res = np.zeros(1, y.dtype)
    l0 = mesh["l0"].astype(y.dtype)
    getattr(lib(), "orc_implicit_pass_" + s)(len(fixed), len(l0), P(mesh["edges"]), P(l0),
                                              P(np.ascontiguousarray(ya, y.dtype)), P(y), P(fixed), P(m_inv),
                                              ct(kappa), P(gs), P(ks), P(res))
    return float(res[0])


def gs_pass(mesh, x, y, fixed, m_inv, k_s, k_b, order, offsets):
    ct, s = _real(x.dtype)
    dt = x.dtype
    dx = np.zeros_like(x)
    nc = np.zeros(len(fixed), dt)
    l0 = mesh["l0"].astype(dt)
    bl0 = mesh["bl0"].astype(dt)
    getattr(lib(), "orc_gs_pass_" + s)(
        len(fixed), len(bl0), P(mesh["bend"]), P(bl0), P(x), P(y), P(fixed), P(m_inv), ct(k_b),
        P(np.ascontiguousarray(order, np.int32)), P(mesh["edges"]), P(l0),
        P(np.ascontiguousarray(offsets, np.int32)), len(offsets) - 1, ct(k_s), P(dx), P(nc))
    return y


def update_vx(x, v, y, fixed, damping, dt, par=False):
    ct, s = _real(x.dtype, par)
    getattr(lib(), "orc_update_vx_" + s)(len(fixed), P(x), P(v), P(y), P(fixed), ct(damping), ct(dt))


def sewing_pass(y, pairs, fixed, m_inv, k):
    """Sewing constraints as one Jacobi pass (XPBD.py:418-441); y in place."""
    ct, s = _real(y.dtype)
    pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    dx = np.zeros_like(y)
    nc = np.zeros(len(fixed), y.dtype)
    getattr(lib(), "orc_sewing_pass_" + s)(len(fixed), len(pr), P(pr), P(y), P(fixed), P(m_inv), ct(k), P(dx), P(nc))
    return y


def compute_velocity(x, v, y, fixed, damping, dt, par=False):
    ct, s = _real(x.dtype, par)
    getattr(lib(), "orc_compute_velocity_" + s)(len(fixed), P(x), P(v), P(y), P(fixed), ct(damping), ct(dt))


def update_x(x, v, dt, par=False):
    ct, s = _real(x.dtype, par)
    getattr(lib(), "orc_update_x_" + s)(len(x), P(x), P(v), ct(dt))


def d_type_pt(p, t0, t1, t2, dtype=np.float32):
    s = "f" if np.dtype(dtype) == np.float32 else "d"
    a = [np.ascontiguousarray(q, dtype) for q in (p, t0, t1, t2)]
    return getattr(lib(), "orc_d_type_pt_" + s)(*[P(q) for q in a])


def dist(dt_, p, t0, t1, t2, dtype=np.float32):
    s = "f" if np.dtype(dtype) == np.float32 else "d"
    a = [np.ascontiguousarray(q, dtype) for q in (p, t0, t1, t2)]
    g = np.zeros(12, dtype)
    d = getattr(lib(), "orc_dist_" + s)(int(dt_), *[P(q) for q in a], P(g))
    return d, g.reshape(4, 3)


def d_type_ee(x0, x1, x2, x3, dtype=np.float32):
    s = "f" if np.dtype(dtype) == np.float32 else "d"
    a = [np.ascontiguousarray(q, dtype) for q in (x0, x1, x2, x3)]
    return getattr(lib(), "orc_d_type_ee_" + s)(*[P(q) for q in a])


def dist_ee(x0, x1, x2, x3, dtype=np.float32):
    s = "f" if np.dtype(dtype) == np.float32 else "d"
    a = [np.ascontiguousarray(q, dtype) for q in (x0, x1, x2, x3)]
    g = np.zeros(12, dtype)
    d = getattr(lib(), "orc_dist_ee_" + s)(*[P(q) for q in a], P(g))
    return d, g.reshape(4, 3)


class _EE(C.Structure):
    _fields_ = [("ne", C.c_int), ("edges_dy", C.c_void_p), ("edges_st", C.c_void_p), ("fe_dy", C.c_void_p),
                ("fe_st", C.c_void_p), ("cap", C.c_int), ("st_c", C.c_void_p), ("st_n", C.c_void_p),
                ("dy_c", C.c_void_p), ("dy_n", C.c_void_p), ("pcap", C.c_int), ("pn_st", C.c_void_p),
                ("pn_dy", C.c_void_p), ("pi_st", C.c_void_p), ("pg_st", C.c_void_p), ("ps_st", C.c_void_p),
                ("pi_dy", C.c_void_p), ("pg_dy", C.c_void_p), ("ps_dy", C.c_void_p)]


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, np.int32)


def _ee_struct(ee, keep, real=None, cap=0, record=False):
    """Fill an orc_ee from a dict: edges_dy, edges_st, fe_dy, fe_st, optional
    st=(cand, count) / dy=(cand, count) candidate faces, pcap; with record, fresh
    pair arrays (returned in `keep`, which also holds every array the struct
    points to).  For the velocity pass, ee also carries the pair arrays."""
    E = _EE()
    ne = len(ee["edges_dy"])
    E.ne = ne
    for k in ("edges_dy", "edges_st", "fe_dy", "fe_st"):
        a = _i32(ee.get(k))
        keep[k] = a
        setattr(E, k, P(a))
    E.cap = cap
    E.pcap = int(ee.get("pcap", 3 * cap))
    for nm in ("st", "dy"):
        cn = ee.get(nm)
        if cn is not None and "pn_" + nm not in ee:
            c, n = _i32(cn[0]), _i32(cn[1])
            keep[nm + "_c"], keep[nm + "_n"] = c, n
            setattr(E, nm + "_c", P(c))
            setattr(E, nm + "_n", P(n))
        if record:
            if cn is None:
                continue
            n1 = max(ne, 1)
            keep["pn_" + nm] = np.zeros(n1, np.int32)
            keep["pi_" + nm] = np.zeros((n1, max(E.pcap, 1), 2), np.int32)
            keep["pg_" + nm] = np.zeros((n1, max(E.pcap, 1), 4, 3), real)
            keep["ps_" + nm] = np.zeros((n1, max(E.pcap, 1)), real)
        elif "pn_" + nm in ee and ee["pn_" + nm] is not None:
            keep["pn_" + nm] = _i32(ee["pn_" + nm])
            keep["pi_" + nm] = _i32(ee["pi_" + nm])
            keep["pg_" + nm] = np.ascontiguousarray(ee["pg_" + nm], real)
            keep["ps_" + nm] = np.ascontiguousarray(ee["ps_" + nm], real)
        else:
            continue
        for f in ("pn_", "pi_", "pg_", "ps_"):
            setattr(E, f + nm, P(keep[f + nm]))
    return E


def morton3d(x, y, z):
    return lib().orc_morton3d(x, y, z)


def expand_bits(v):
    return lib().orc_expand_bits(v)


def cell_morton(lo, hi):
    codes = np.zeros(NCELLS, np.int32)
    centers = np.zeros((NCELLS, 3), np.float32)
    cs = np.zeros(3, np.float32)
    lib().orc_cell_morton(P(np.asarray(lo, np.float32)), P(np.asarray(hi, np.float32)),
                          P(codes), P(centers), P(cs))
    return codes, centers, cs


def radix_sort(keys, ids, bits, passes):
    k = np.ascontiguousarray(keys, np.int32).copy()
    i = np.ascontiguousarray(ids, np.int32).copy()
    lib().orc_radix_sort(len(k), P(k), P(i), bits, passes)
    return k, i


def karras(sorted_codes, wide):
    n = len(sorted_codes)
    sc = np.ascontiguousarray(sorted_codes, np.int32)
    ca = np.full(max(n - 1, 1), -1, np.int32)
    cb = np.full(max(n - 1, 1), -1, np.int32)
    par = np.full(max(2 * n - 1, 1), -1, np.int32)
    rl = np.zeros(max(n - 1, 1), np.int32)
    rr = np.zeros(max(n - 1, 1), np.int32)
    lib().orc_karras(n, P(sc), int(wide), P(ca), P(cb), P(par), P(rl), P(rr))
    return dict(child_a=ca[:n - 1], child_b=cb[:n - 1], parent=par[:2 * n - 1],
                range_l=rl[:n - 1], range_r=rr[:n - 1])


def face_aabbs(faces, x, padding):
    nf = len(faces)
    fmin = np.zeros((nf, 3), np.float32)
    fmax = np.zeros((nf, 3), np.float32)
    lib().orc_face_aabbs(nf, P(np.ascontiguousarray(faces, np.int32)), P(np.ascontiguousarray(x, np.float32)),
                         C.c_float(padding), P(fmin), P(fmax))
    return fmin, fmax


def mesh_aabb(x, padding):
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().orc_mesh_aabb(len(x), P(np.ascontiguousarray(x, np.float32)), C.c_float(padding), P(lo), P(hi))
    return lo, hi


class CellTopology:
    """Static 32^3 cell tree built once (lbvh_cell.py:78-80)."""

    def __init__(self):
        codes, _, _ = cell_morton(np.zeros(3, np.float32), np.full(3, 1e3, np.float32))
        self.codes = codes
        ids = np.arange(NCELLS, dtype=np.int32)
        self.sorted_codes, self.sorted_ids = radix_sort(codes, ids, 6, 5)
        t = karras(self.sorted_codes, wide=False)
        self.child_a, self.child_b, self.parent = t["child_a"], t["child_b"], t["parent"]
        self.range_l, self.range_r = t["range_l"], t["range_r"]


_CELL_TOPO = None


def cell_topology():
    global _CELL_TOPO
    if _CELL_TOPO is None:
        _CELL_TOPO = CellTopology()
    return _CELL_TOPO


class CellBVH:
    """LBVH_CELL.build (lbvh_cell.py:637-653) over face boxes of `x`."""

    def __init__(self, faces, x, padding):
        topo = cell_topology()
        self.kind = 0
        self.fmin, self.fmax = face_aabbs(faces, x, padding)
        self.lo, self.hi = mesh_aabb(x, padding)
        nf = len(faces)
        nn = 2 * NCELLS - 1
        self.cell_size = np.zeros(3, np.float32)
        self.face_cell = np.zeros(nf, np.int32)
        self.prefix = np.zeros(NCELLS, np.int32)
        self.sorted_faces = np.zeros(max(nf, 1), np.int32)
        self.num = np.zeros(nn, np.int32)
        self.leaf_l = np.zeros(NCELLS, np.int32)
        self.leaf_r = np.zeros(NCELLS, np.int32)
        self.nmin = np.zeros((nn, 3), np.float32)
        self.nmax = np.zeros((nn, 3), np.float32)
        self.child_a, self.child_b = topo.child_a, topo.child_b
        self.bad = lib().orc_cell_build(
            nf, P(self.fmin), P(self.fmax), P(self.lo), P(self.hi), P(topo.child_a), P(topo.child_b),
            P(topo.parent), P(self.cell_size), P(self.face_cell), P(self.prefix), P(self.sorted_faces),
            P(self.num), P(self.leaf_l), P(self.leaf_r), P(self.nmin), P(self.nmax))
        self.n_leaves = NCELLS
        self.origin = self.lo

    def query(self, qmin, qmax, cap):
        return _traverse(self, qmin, qmax, cap)


class TriBVH:
    """Per-triangle LBVH: Morton sort of face-box centres (normalised by the cube
    of the mesh box's largest extent), Karras with index tie-break, refit."""

    def __init__(self, faces, x, padding):
        self.kind = 1
        self.faces = np.ascontiguousarray(faces, np.int32)
        self.padding = padding
        self.fmin, self.fmax = face_aabbs(faces, x, padding)
        self.lo, self.hi = mesh_aabb(x, padding)
        nf = len(faces)
        self.n_leaves = nf
        self.codes = np.zeros(nf, np.int32)
        self.sorted_faces = np.zeros(nf, np.int32)
        self.child_a = np.zeros(max(nf - 1, 1), np.int32)
        self.child_b = np.zeros(max(nf - 1, 1), np.int32)
        self.parent = np.zeros(max(2 * nf - 1, 1), np.int32)
        self.nmin = np.zeros((max(2 * nf - 1, 1), 3), np.float32)
        self.nmax = np.zeros((max(2 * nf - 1, 1), 3), np.float32)
        lib().orc_tri_build(nf, P(self.fmin), P(self.fmax), P(self.lo), P(self.hi), P(self.codes),
                            P(self.sorted_faces), P(self.child_a), P(self.child_b), P(self.parent),
                            P(self.nmin), P(self.nmax))
        self.num = np.ones(max(2 * nf - 1, 1), np.int32)
        self.cell_size = np.ones(3, np.float32)
        self.origin = np.zeros(3, np.float32)

    def refit(self, x):
        self.fmin, self.fmax = face_aabbs(self.faces, x, self.padding)
        lib().orc_tri_refit(self.n_leaves, P(self.fmin), P(self.fmax), P(self.sorted_faces),
                            P(self.child_a), P(self.child_b), P(self.nmin), P(self.nmax))

    def query(self, qmin, qmax, cap):
        return _traverse(self, qmin, qmax, cap)


def _traverse(t, qmin, qmax, cap):
    nq = len(qmin)
    cand = np.full((max(nq, 1), cap), -1, np.int32)
    cnt = np.zeros(max(nq, 1), np.int32)
    dummy = np.zeros(1, np.int32)
    flagged = lib().orc_traverse(
        t.kind, nq, P(np.ascontiguousarray(qmin, np.float32)), P(np.ascontiguousarray(qmax, np.float32)),
        t.n_leaves, P(t.child_a), P(t.child_b), P(t.num),
        P(getattr(t, "leaf_l", dummy)), P(getattr(t, "leaf_r", dummy)), P(t.sorted_faces),
        P(t.nmin), P(t.nmax), P(t.fmin), P(t.fmax), P(t.cell_size), P(t.origin), cap, P(cand), P(cnt))
    return cand[:nq], cnt[:nq], flagged


def greedy_coloring(nv, edges, max_colors=64):
    e = np.ascontiguousarray(edges, np.int32)
    ne = len(e)
    color = np.zeros(ne, np.int32)
    order = np.zeros(ne, np.int32)
    offsets = np.zeros(max_colors + 1, np.int32)
    nc = lib().orc_greedy_coloring(nv, ne, P(e), max_colors, P(color), P(order), P(offsets))
    if nc < 0:
        raise RuntimeError("more than max_colors colours needed")
    return color, order, offsets[:nc + 1]


def collision_pass(mesh_dy, faces_st, y, xs, fixed, m_inv, dhat, k_col, cap, vt_st, vt_dy, tv_st, record=False,
                   ee=None, rec_bufs=None, par=False):
    """Position-level collision pass; with record=True the result also holds the
    pair records of the three query sets ("pairs": per set a dict of num [n],
    fid/dtype [n, cap], g [n, cap, 4, 3], schur [n, cap]), the reference's
    vt_st_pair/vt_dy_pair/tv_st_pair fields (XPBD.py:51-73).

    ee: edge-edge sets (XPBD.py:266-281), a dict of edges_dy [ne, 2], edges_st,
    fe_dy [F, 3], fe_st (face -> edge ids), st / dy = (cand [ne, cap], count) or
    None, pcap (pair records kept per edge, default 3 cap).  The result's "ee"
    holds per set the contacting pair count pn_* (uncapped) and the records
    pi_* (other edge, type), pg_*, ps_* (with record=True), ready for
    velocity_pass.

    rec_bufs: a list the caller keeps between passes; the record arrays are
    allocated into it once and reused (records past a query's count then hold
    an earlier pass's values, which nothing reads), instead of ~60 bytes x cap
    of fresh zeroed memory per query and pass."""
    ct, s = _real(y.dtype, par)
    nv = len(fixed)
    nvs = 0 if xs is None else len(xs)
    dx = np.zeros_like(y)
    nc = np.zeros(nv, y.dtype)
    pn = [np.zeros(max(nv, 1), np.int32), np.zeros(max(nv, 1), np.int32), np.zeros(max(nvs, 1), np.int32)]
    rec = []
    for n in (nv, nv, nvs):
        n = max(n, 1)
        rec.append((np.zeros((n, cap, 2), np.int32), np.zeros((n, cap, 4, 3), y.dtype),
                    np.zeros((n, cap), y.dtype)) if record and rec_bufs is None else (None, None, None))
    if record and rec_bufs is not None:
        shapes = [((max(n, 1), cap, 2), np.int32) for n in (nv, nv, nvs)]
        if not rec_bufs or any(b[0].shape != s[0] or b[1].dtype != y.dtype for b, s in zip(rec_bufs, shapes)):
            rec_bufs[:] = [(np.zeros((max(n, 1), cap, 2), np.int32), np.zeros((max(n, 1), cap, 4, 3), y.dtype),
                            np.zeros((max(n, 1), cap), y.dtype)) for n in (nv, nv, nvs)]
        rec = list(rec_bufs)

    def cn(t):
        return (None, None) if t is None else (np.ascontiguousarray(t[0], np.int32), np.ascontiguousarray(t[1], np.int32))
    (a, an), (b, bn), (c, cnn) = cn(vt_st), cn(vt_dy), cn(tv_st)
    xs_ = np.zeros((1, 3), y.dtype) if xs is None else np.ascontiguousarray(xs, y.dtype)
    fst = np.zeros((1, 3), np.int32) if faces_st is None else np.ascontiguousarray(faces_st, np.int32)
    args = (nv, nvs, P(mesh_dy["faces"]), P(fst), P(y), P(xs_), P(fixed), P(m_inv), ct(dhat), ct(k_col), cap,
            P(a), P(an), P(b), P(bn), P(c), P(cnn), P(dx), P(nc), P(pn[0]), P(pn[1]), P(pn[2]),
            *[P(arr) for r in rec for arr in r])
    out = dict(dx=dx, nc=nc, pair_num=pn)
    if ee is None:
        getattr(lib(), "orc_collision_pass_rec_" + s)(*args)
    else:
        keep = {}
        E = _ee_struct(ee, keep, real=y.dtype, cap=cap, record=True)
        if not record:
            for nm in ("st", "dy"):
                for f in ("pi_", "pg_", "ps_"):
                    setattr(E, f + nm, None)
                    keep.pop(f + nm, None)
        getattr(lib(), "orc_collision_pass_ee_" + s)(*args, C.byref(E))
        out["ee"] = dict(edges_dy=keep["edges_dy"], edges_st=keep["edges_st"], pcap=E.pcap,
                         **{k: keep.get(k) for nm in ("st", "dy") for k in
                            ("pn_" + nm, "pi_" + nm, "pg_" + nm, "ps_" + nm)})
    if record:
        present = (vt_st is not None, vt_dy is not None, tv_st is not None)
        out["pairs"] = [dict(num=pn[k], fid=rec[k][0][..., 0], dtype=rec[k][0][..., 1], g=rec[k][1],
                             schur=rec[k][2], pi=rec[k][0]) if present[k] else None for k in range(3)]
        out["cap"] = cap
    return out


def velocity_pass(mesh_dy, faces_st, v, vs, fixed, m_inv, m_inv_st, mu, pairs, cap, nvs=None, ee=None, par=False):
    """solve_constraints_v (XPBD.py:374-381, 283-311; collision_constraints_v.py
    5-570; update_dv 397-404) over pair records from collision_pass(record=True);
    v is updated in place.  pairs None: no collision, update_dv alone.
    vs None: static vertices at rest.  nvs: static vertex count (default: from
    vs, m_inv_st, the static-vertex pair records or the static faces).
    ee: collision_pass(record=True, ee=...)["ee"] -- the edge-edge pairs follow
    the vertex-triangle ones (XPBD.py:313-330; collision_constraints_v.py
    573-878)."""
    ct, s = _real(v.dtype, par)
    nv = len(fixed)
    if nvs is None:
        if vs is not None:
            nvs = len(vs)
        elif m_inv_st is not None:
            nvs = len(m_inv_st)
        elif pairs is not None and pairs[2] is not None:
            nvs = len(pairs[2]["num"])
        elif faces_st is not None:
            nvs = int(np.max(faces_st)) + 1
        else:
            nvs = 0
    dv = np.zeros_like(v)
    nc = np.zeros(nv, v.dtype)
    vs_ = np.zeros((max(nvs, 1), 3), v.dtype) if vs is None else np.ascontiguousarray(vs, v.dtype)
    fst = np.zeros((1, 3), np.int32) if faces_st is None else np.ascontiguousarray(faces_st, np.int32)
    mst = None if m_inv_st is None else np.ascontiguousarray(m_inv_st, v.dtype)
    args = []
    for k in range(3):
        p = None if pairs is None else pairs[k]
        args += [None, None, None, None] if p is None else [P(p["num"]), P(p["pi"]), P(p["g"]), P(p["schur"])]
    call = (nv, nvs, P(mesh_dy["faces"]), P(fst), P(v), P(vs_), P(fixed), P(m_inv), P(mst), ct(mu), cap, *args,
            P(dv), P(nc))
    if ee is None:
        getattr(lib(), "orc_velocity_pass_" + s)(*call)
    else:
        keep = {}
        E = _ee_struct(ee, keep, real=v.dtype)
        getattr(lib(), "orc_velocity_pass_ee_" + s)(*call, C.byref(E))
    return dict(dv=dv, nc=nc)


# ---------------------------------------------------------------- full driver

class OracleCloth:
    """Serial restatement of Solver.forward (XPBD.py:443-469).

    mesh_dy / mesh_st: dicts with x (V,3) float32, faces (F,3) int32; mesh_dy also
    edges, l0, bend, bl0 (from the package's topology, identical to the GPU's
    inputs), fixed (V,), m_inv (V,).
    """

    def __init__(self, mesh_dy, mesh_st=None, *, dhat, stiffness_stretch, stiffness_bending,
                 g, dt, dtype=np.float32, damping=0.001, padding=0.05, collision=False,
                 bvh="cell", cap=40, refit_every_substep=False, method="jacobi", velocity=False, mu=0.8,
                 sewing=None, sewing_stiffness=None, edge_edge=False, ee_pcap=None, parallel=False):
        """edge_edge: the edge-edge constraints of XPBD.py:217-232, 266-281,
        313-330 (mesh_dy and mesh_st then also carry edges [E, 2] and fe [F, 3],
        the edge ids of each face's sides); ee_pcap: edge pairs kept per edge for
        the velocity pass (default cap).  parallel: run the element loops on
        OpenMP threads with atomic scatter (float32; a timing baseline, not a
        checker: sums are then in no fixed order)."""
        self.par = bool(parallel)
        self.method = method
        self.edge_edge, self.ee_pcap = edge_edge, (cap if ee_pcap is None else ee_pcap)
        self.velocity, self.mu = velocity, mu
        self.sewing = None if sewing is None else np.ascontiguousarray(sewing, np.int32).reshape(-1, 2)
        self.sewing_stiffness = sewing_stiffness
        self.dt_ = np.dtype(dtype)
        self.m = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in mesh_dy.items()}
        self.m["edges"] = np.ascontiguousarray(self.m["edges"], np.int32)
        self.m["bend"] = np.ascontiguousarray(self.m["bend"], np.int32)
        self.m["faces"] = np.ascontiguousarray(self.m["faces"], np.int32)
        self.st = mesh_st
        self.x = np.ascontiguousarray(mesh_dy["x"], self.dt_).copy()
        self.v = np.zeros_like(self.x)
        self.y = self.x.copy()
        self.fixed = np.ascontiguousarray(mesh_dy["fixed"], self.dt_).copy()
        self.m_inv = np.ascontiguousarray(mesh_dy["m_inv"], self.dt_).copy()
        if self.dt_ == np.float64:
            # FP64 oracle: rest lengths recomputed in double from the same positions
            x0 = np.asarray(mesh_dy["x0"], np.float64)
            e, b = self.m["edges"], self.m["bend"]
            self.m["l0"] = np.linalg.norm(x0[e[:, 0]] - x0[e[:, 1]], axis=1)
            self.m["bl0"] = np.linalg.norm(x0[b[:, 0]] - x0[b[:, 1]], axis=1)
        self.dhat, self.ks, self.kb = dhat, stiffness_stretch, stiffness_bending
        self.g = np.asarray(g, np.float32)
        self.dt, self.damping, self.padding = dt, damping, padding
        self.collision, self.bvh, self.cap = collision, bvh, cap
        self.refit_every_substep = refit_every_substep
        self.stats = dict(overflow=0)
        self._rec_bufs = []

    def _trees(self):
        xf = self.x.astype(np.float32)
        Tree = CellBVH if self.bvh == "cell" else TriBVH
        self.tree_dy = Tree(self.m["faces"], xf, self.padding)
        self.tree_st = Tree(self.st["faces"], self.st["x"], self.padding) if self.st is not None else None

    def _broadphase(self):
        pad = np.float32(self.padding)
        yf = self.y.astype(np.float32)
        qmin, qmax = yf - pad, yf + pad
        out = {}
        if self.tree_st is not None:
            c, n, fl = self.tree_st.query(qmin, qmax, self.cap)
            out["vt_st"] = (c, n)
            self.stats["overflow"] += fl
            xs = np.asarray(self.st["x"], np.float32)
            c, n, fl = self.tree_dy.query(xs - pad, xs + pad, self.cap)
            out["tv_st"] = (c, n)
            self.stats["overflow"] += fl
        c, n, fl = self.tree_dy.query(qmin, qmax, self.cap)
        out["vt_dy"] = (c, n)
        self.stats["overflow"] += fl
        if self.edge_edge:
            # edge boxes: min/max of the end points -+ padding (XPBD.py:217-232)
            e = self.m["edges"]
            emin = np.minimum(yf[e[:, 0]], yf[e[:, 1]]) - pad
            emax = np.maximum(yf[e[:, 0]], yf[e[:, 1]]) + pad
            if self.tree_st is not None:
                c, n, fl = self.tree_st.query(emin, emax, self.cap)
                out["ee_st"] = (c, n)
                self.stats["overflow"] += fl
            c, n, fl = self.tree_dy.query(emin, emax, self.cap)
            out["ee_dy"] = (c, n)
            self.stats["overflow"] += fl
        return out

    def forward(self, n_substeps, n_iters=1):
        dt_sub = self.dt / n_substeps
        R = self.dt_.type
        k_s = R(self.kb * dt_sub * dt_sub)   # XPBD.py:354-355 (names swapped there)
        k_b = R(self.ks * dt_sub * dt_sub)
        dts = R(np.float32(dt_sub)) if self.dt_ == np.float32 else R(dt_sub)
        g = self.g.astype(self.dt_)
        if self.collision:
            self._trees()
        for _ in range(n_substeps):
            if self.collision and self.bvh == "tri" and self.refit_every_substep:
                self.tree_dy.refit(self.x.astype(np.float32))
            self.y = compute_y(self.x, self.v, self.fixed, dts, g, self.par)
            if self.method == "gs" and not hasattr(self, "_colors"):
                _, order, off = greedy_coloring(len(self.fixed), self.m["edges"])
                self._colors = (order, off)
            for _ in range(n_iters):
                if self.method == "gs":
                    gs_pass(self.m, self.x, self.y, self.fixed, self.m_inv, k_s, k_b, *self._colors)
                else:
                    spring_pass(self.m, self.x, self.y, self.fixed, self.m_inv, k_s, k_b, self.par)
                if self.collision:
                    cands = self._broadphase()
                    xs = None if self.st is None else np.asarray(self.st["x"], self.dt_)
                    ee = None
                    if self.edge_edge:
                        st = self.st
                        ee = dict(edges_dy=self.m["edges"], fe_dy=self.m["fe"],
                                  edges_st=None if st is None else st["edges"],
                                  fe_st=None if st is None else st["fe"],
                                  st=cands.get("ee_st"), dy=cands["ee_dy"], pcap=self.ee_pcap)
                    res = collision_pass(self.m, None if self.st is None else self.st["faces"], self.y, xs,
                                         self.fixed, self.m_inv, R(self.dhat), R(1e8), self.cap,
                                         cands.get("vt_st"), cands.get("vt_dy"), cands.get("tv_st"),
                                         record=self.velocity, ee=ee, rec_bufs=self._rec_bufs, par=self.par)
                    self.last_pairs = res.get("pairs")
                    self.last_ee = res.get("ee")
                if self.sewing is not None and len(self.sewing):
                    ksew = 0.5 * self.kb if self.sewing_stiffness is None else self.sewing_stiffness
                    sewing_pass(self.y, self.sewing, self.fixed, self.m_inv, R(ksew * dt_sub * dt_sub))
            if not self.velocity:
                update_vx(self.x, self.v, self.y, self.fixed, R(np.float32(self.damping)), dts, self.par)
                continue
            # compute_velocity, solve_constraints_v, update_x (XPBD.py:465-469)
            compute_velocity(self.x, self.v, self.y, self.fixed, R(np.float32(self.damping)), dts, self.par)
            st = self.st
            vs = None if st is None else np.ascontiguousarray(st.get("v", np.zeros_like(st["x"])), self.dt_)
            mst = None if st is None or "m_inv" not in st else np.asarray(st["m_inv"], self.dt_)
            velocity_pass(self.m, None if st is None else st["faces"], self.v, vs, self.fixed, self.m_inv, mst,
                          R(np.float32(self.mu)), self.last_pairs if self.collision else None, self.cap,
                          ee=self.last_ee if self.collision and self.edge_edge else None, par=self.par)
            update_x(self.x, self.v, dts, self.par)
        return self.x
