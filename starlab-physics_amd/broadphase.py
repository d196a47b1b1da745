"""Device LBVH over a mesh's triangles: the reference's LBVH_CELL
(framework/collision/lbvh_cell.py:17-80, build 637-653, traversal 661-714) and a
per-triangle LBVH, both backed by libslx.so.

    tree = DeviceTree(mesh, mode="cell")   # or "tri"
    tree.build(padding)                     # from mesh.x4
    cand, n = tree.query(points4, padding, cap)
"""
from __future__ import annotations

import torch

from . import _abi

NCELLS = 32768
_TOPO = {}


def cell_topology(device):
    """Static 32^3 cell tree (lbvh_cell.py:78-80), built once per device."""
    key = str(device)
    if key not in _TOPO:
        i32 = dict(dtype=torch.int32, device=device)
        t = dict(codes=torch.empty(NCELLS, **i32), sorted_ids=torch.empty(NCELLS, **i32),
                 child_a=torch.empty(NCELLS - 1, **i32), child_b=torch.empty(NCELLS - 1, **i32),
                 parent=torch.empty(2 * NCELLS - 1, **i32), range_l=torch.empty(NCELLS - 1, **i32),
                 range_r=torch.empty(NCELLS - 1, **i32),
                 sched=torch.empty(_abi.lib().slx_refit_sched_ints(NCELLS), **i32))
        ws = torch.empty(_abi.lib().slx_cell_topology_ws_bytes(), dtype=torch.uint8, device=device)
        _abi.call("slx_cell_topology", *[_abi.ptr(t[k]) for k in
                                         ("codes", "sorted_ids", "child_a", "child_b", "parent", "range_l",
                                          "range_r", "sched")], _abi.ptr(ws), ws.numel(), _abi.stream_handle())
        _TOPO[key] = t
    return _TOPO[key]


class CandidateCache:
    """Device buffers of one query set's candidate cache (slx_cand_cache)."""

    def __init__(self, nq, cap, margin, device):
        """cap 0: no cache, only the redo list of the two-tier walk."""
        i32 = dict(dtype=torch.int32, device=device)
        self.nq, self.cap, self.margin = nq, cap, margin
        # int4 rows of four leaves per query: [(cap + 3) // 4][nq]
        self.leaves = torch.zeros(max(4 * ((cap + 3) // 4) * nq, 4), **i32) if cap > 0 else None
        self.count = torch.zeros(max(nq, 1), **i32) if cap > 0 else None
        self.q0 = torch.zeros((max(nq, 1), 4), dtype=torch.float32, device=device) if cap > 0 else None
        self.redo = torch.zeros(nq + 1, **i32)

    def struct(self) -> _abi.CandCache:
        P = _abi.ptr
        return _abi.CandCache(int(self.cap), float(self.margin), P(self.leaves), P(self.count), P(self.q0),
                              P(self.redo))

    def redone(self) -> int:
        """Queries of the last pass walked in full (synchronises)."""
        return int(self.redo[0].item())


class DeviceTree:
    def __init__(self, mesh, mode="cell"):
        if mode not in ("cell", "tri"):
            raise ValueError("mode must be 'cell' or 'tri'")
        self.mesh = mesh
        self.mode = mode
        dev = mesh.x4.device
        nf = mesh.num_faces
        self.nf = nf
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.fmin4 = torch.zeros((max(nf, 1), 4), **f32)
        self.fmax4 = torch.zeros((max(nf, 1), 4), **f32)
        self.box6 = torch.zeros(8, **f32)
        self.sorted_faces = torch.zeros(max(nf, 1), **i32)
        self.faces = mesh.face_indices
        if mode == "cell":
            topo = cell_topology(dev)
            self.child_a, self.child_b, self.parent = topo["child_a"], topo["child_b"], topo["parent"]
            self.range_l, self.range_r, self.sched = topo["range_l"], topo["range_r"], topo["sched"]
            nn = 2 * NCELLS - 1
            self.num = torch.zeros(nn, **i32)
            self.face_cell = torch.zeros(max(nf, 1), **i32)
            self.prefix = torch.zeros(NCELLS, **i32)
            self.leaf_l = torch.zeros(NCELLS, **i32)
            self.leaf_r = torch.zeros(NCELLS, **i32)
            self.codes = None
            self.n_leaves = NCELLS
            ws = _abi.lib().slx_cell_build_ws_bytes(nf)
        else:
            nn = max(2 * nf - 1, 1)
            self.child_a = torch.zeros(max(nf - 1, 1), **i32)
            self.child_b = torch.zeros(max(nf - 1, 1), **i32)
            self.parent = torch.zeros(nn, **i32)
            self.codes = torch.zeros(max(nf, 1), **i32)
            self.range_l = torch.zeros(max(nf - 1, 1), **i32)
            self.range_r = torch.zeros(max(nf - 1, 1), **i32)
            self.sched = torch.zeros(_abi.lib().slx_refit_sched_ints(max(nf, 1)), **i32)
            self.num = self.face_cell = self.prefix = self.leaf_l = self.leaf_r = None
            self.n_leaves = nf
            ws = _abi.lib().slx_tri_build_ws_bytes(nf)
        self.nlo4 = torch.zeros((nn, 4), **f32)
        self.nhi4 = torch.zeros((nn, 4), **f32)
        # triangle trees: 4-wide node array (grandchild boxes of every internal node)
        self.quad4 = torch.zeros((max(nf - 1, 1) * 8, 4), **f32) if mode == "tri" else None
        # batched instances: one subtree per instance (slx_seg_build)
        self.inst = getattr(mesh, "inst", None)
        self.n_inst = 0
        if self.inst is not None and self.inst["n"] > 1:
            if mode != "tri":
                raise ValueError("batched instances need the triangle LBVH (mode='tri')")
            ni = self.inst["n"]
            self.n_inst = ni
            self.box6 = torch.zeros(6 * ni + 2, **f32)
            self.sorted_inst = torch.zeros(max(nf, 1), **i32)
            self.roots = torch.zeros(ni, **i32)
            ws = max(ws, _abi.lib().slx_seg_build_ws_bytes(nf, ni))
        ws = max(ws, _abi.lib().slx_mesh_aabb_ws_bytes())
        self.ws = torch.empty(ws, dtype=torch.uint8, device=dev)
        self.status = torch.zeros(1, **i32)
        # candidate caches over this tree: vertices at the build and their displacement since
        self.x_snap = None
        self.disp = None

    def enable_snapshot(self):
        if self.x_snap is None:
            self.x_snap = torch.zeros((max(self.mesh.num_verts, 1), 4), dtype=torch.float32,
                                      device=self.mesh.x4.device)
            self.disp = torch.zeros(1, dtype=torch.int32, device=self.mesh.x4.device)

    def struct(self) -> _abi.Tree:
        P = _abi.ptr
        t = _abi.Tree(0 if self.mode == "cell" else 1, self.nf, P(self.fmin4), P(self.fmax4), P(self.box6),
                      P(self.codes), P(self.sorted_faces), P(self.child_a), P(self.child_b), P(self.parent),
                      P(self.num), P(self.face_cell), P(self.prefix), P(self.leaf_l), P(self.leaf_r),
                      P(self.nlo4), P(self.nhi4))
        t.range_l, t.range_r, t.sched = P(self.range_l), P(self.range_r), P(self.sched)
        t.quad4 = P(self.quad4)
        t.x_snap, t.disp = P(self.x_snap), P(self.disp)
        t.vert_order = P(getattr(self.mesh, "vert_order", None))
        if self.n_inst > 1:
            I = self.inst
            t.n_inst = self.n_inst
            t.vert_base, t.face_base = P(I["vert_base"]), P(I["face_base"])
            t.face_inst, t.vert_inst = P(I["face_inst"]), P(I["vert_inst"])
            t.sorted_inst, t.roots = P(self.sorted_inst), P(self.roots)
        return t

    # --- stages (the frame driver calls the same entry points itself) ---
    def face_aabbs(self, padding, x4=None):
        x4 = self.mesh.x4 if x4 is None else x4
        _abi.call("slx_face_aabbs", self.nf, _abi.ptr(self.faces), _abi.ptr(x4), float(padding),
                  _abi.ptr(self.fmin4), _abi.ptr(self.fmax4), _abi.stream_handle())

    def mesh_aabb(self, padding, x4=None):
        x4 = self.mesh.x4 if x4 is None else x4
        _abi.call("slx_mesh_aabb", self.mesh.num_verts, _abi.ptr(x4), float(padding), _abi.ptr(self.box6),
                  _abi.ptr(self.ws), self.ws.numel(), _abi.stream_handle())

    def build(self, padding, x4=None):
        """computeAABB_faces + computeAABB + LBVH build from x (XPBD.py:450-452)."""
        P = _abi.ptr
        s = _abi.stream_handle()
        if self.n_inst > 1:
            x4 = self.mesh.x4 if x4 is None else x4
            I = self.inst
            _abi.call("slx_seg_build", self.mesh.num_verts, self.nf, self.n_inst, P(self.faces), P(x4),
                      float(padding), P(I["vert_base"]), P(I["face_base"]), P(I["face_inst"]), P(self.fmin4),
                      P(self.fmax4), P(self.box6), P(self.codes), P(self.sorted_faces), P(self.sorted_inst),
                      P(self.child_a), P(self.child_b), P(self.parent), P(self.range_l), P(self.range_r),
                      P(self.roots), P(self.sched), P(self.nlo4), P(self.nhi4), P(self.quad4), P(self.ws),
                      self.ws.numel(), s)
            return
        self.face_aabbs(padding, x4)
        self.mesh_aabb(padding, x4)
        if self.mode == "cell":
            _abi.call("slx_cell_build", self.nf, P(self.fmin4), P(self.fmax4), P(self.box6), P(self.child_a),
                      P(self.child_b), P(self.parent), P(self.sched), P(self.face_cell), P(self.prefix), P(self.sorted_faces),
                      P(self.num), P(self.leaf_l), P(self.leaf_r), P(self.nlo4), P(self.nhi4), P(self.status),
                      P(self.ws), self.ws.numel(), s)
        else:
            _abi.call("slx_tri_build", self.nf, P(self.fmin4), P(self.fmax4), P(self.box6), P(self.codes),
                      P(self.sorted_faces), P(self.child_a), P(self.child_b), P(self.parent), P(self.range_l),
                      P(self.range_r), P(self.sched), P(self.nlo4), P(self.nhi4), P(self.quad4), P(self.ws),
                      self.ws.numel(), s)

    def refit(self, padding, x4=None):
        """Triangle tree: node boxes from new positions, same topology (the face
        boxes fmin4/fmax4 keep the values of the last build)."""
        if self.mode != "tri":
            raise ValueError("refit needs the triangle LBVH")
        P = _abi.ptr
        x4 = self.mesh.x4 if x4 is None else x4
        _abi.call("slx_tri_refit", self.nf, P(self.faces), P(x4), float(padding), P(self.sorted_faces),
                  P(self.child_a), P(self.child_b), P(self.parent), P(self.sched), P(self.nlo4), P(self.nhi4),
                  P(self.quad4), _abi.stream_handle())

    def query(self, q4, padding, cap, cand=None, cand_n=None, status=None, q_inst=None):
        """Candidate faces of boxes q4.xyz -+ padding, column-major (cap, nq)."""
        nq = q4.shape[0]
        dev = q4.device
        if cand is None:
            cand = torch.full((max(cap, 1), max(nq, 1)), -1, dtype=torch.int32, device=dev)
        if cand_n is None:
            cand_n = torch.zeros(max(nq, 1), dtype=torch.int32, device=dev)
        st = self.status if status is None else status
        P = _abi.ptr
        if self.n_inst > 1:
            if q_inst is None:
                raise ValueError("batched tree queries need q_inst (instance of each query)")
            _abi.call("slx_traverse_seg", nq, P(q4), P(q_inst), P(self.roots), float(padding), self.nf,
                      P(self.nlo4), P(self.nhi4), P(self.quad4), cap, P(cand), P(cand_n), P(st),
                      _abi.stream_handle())
            return cand, cand_n
        _abi.call("slx_traverse", 0 if self.mode == "cell" else 1, nq, P(q4), float(padding),
                  self.n_leaves if self.nf > 0 else 0, P(self.num), P(self.sorted_faces), P(self.nlo4),
                  P(self.nhi4), P(self.quad4), P(self.fmin4), P(self.fmax4), P(self.box6), cap, P(cand), P(cand_n),
                  P(st), _abi.stream_handle())
        return cand, cand_n

    def query_edges(self, edges, x4, padding, cap, cand=None, cand_n=None, status=None, order=None, redo=None):
        """Candidate faces of the edge boxes min/max(x4[e0], x4[e1]) -+ padding
        (edges: [ne*2] int32 device), column-major (cap, ne) -- the edge broad
        phase of XPBD.py:217-232.  order: walk order of the edges; redo ([1 + ne]
        int32, triangle trees): the two-tier walk.  Same lists either way."""
        ne = edges.numel() // 2
        dev = x4.device
        if self.n_inst > 1:
            raise ValueError("edge queries of batched trees are not supported")
        if cand is None:
            cand = torch.full((max(cap, 1), max(ne, 1)), -1, dtype=torch.int32, device=dev)
        if cand_n is None:
            cand_n = torch.zeros(max(ne, 1), dtype=torch.int32, device=dev)
        st = self.status if status is None else status
        P = _abi.ptr
        _abi.call("slx_traverse_edges", 0 if self.mode == "cell" else 1, ne, P(edges), P(x4), float(padding),
                  self.n_leaves if self.nf > 0 else 0, P(self.num), P(self.sorted_faces), P(self.nlo4),
                  P(self.nhi4), P(self.quad4), P(self.fmin4), P(self.fmax4), P(self.box6), cap, P(cand), P(cand_n),
                  P(st), P(order), P(redo), _abi.stream_handle())
        return cand, cand_n
