"""Solver: the reference's XPBD.Solver surface over the gfx950 kernels.

Mirrors framework/physics/XPBD.py:9-106 (constructor and attributes),
forward (443-469), reset (108-112) and set_fixed_vertices (389-395).

Extensions (all default to the reference's behaviour):
  forward(n_substeps, n_iters=1)   n_iters spring passes per substep (the
                                   reference runs one, XPBD.py:463)
  bvh_mode = "cell" | "tri"        the reference's 32^3 cell LBVH, or one leaf
                                   per triangle (needed for large cloths)
  refit_every_substep              triangle tree: refit from x at each substep
  cache_size (40)                  candidate capacity per query (XPBD.py:51)
  pair_cache_size (None)           contacting pairs kept per query for the
                                   velocity pass (None = cache_size, which never
                                   drops one; a smaller value saves memory)
  candidate_cache (True)           triangle trees: each query set's candidate
                                   lists come from a per-frame cache built with
                                   the query box enlarged by cache_margin (None =
                                   padding / 4), filtered by the exact box each
                                   pass; the lists are the traversal's, bit for bit
  cache_capacity (None)            leaves stored per query (None = 2 cache_size)
  cache_every (1)                  substeps between cache rebuilds (<= 0: once
                                   per frame)
  static_cache_every (4)           the same for the cloth vertices' cache of the
  static_cache_margin (None)       static tree (None: cache_every), and its margin
                                   (None = 2 x padding): that tree never moves, so
                                   an entry stays valid while its query stays
                                   within the margin
  load_sewing_pairs(path | pairs)  sewing pairs (XPBD.py:405-417); with
  enable_sewing (False)            they are applied after the collision pass of
                                   every iteration (XPBD.py:418-441, whose call
                                   the reference leaves commented out), stiffness
                                   sewing_stiffness (None = 0.5 stiffness_bending,
                                   the reference's "0.5 YM")
  enable_edge_edge (False)         edge-edge constraints (XPBD.py:217-232, 266-281,
                                   313-330, commented out in the reference): each
                                   collision pass also queries the dynamic edges'
                                   boxes and applies __ee_st / __ee_dy after the
                                   vertex-triangle constraints; with the velocity
                                   pass, their velocity-level counterparts
  ee_pair_cache_size (None)        edge pairs kept per edge and set for the
                                   velocity pass (None = cache_size, the
                                   reference's 40-wide fields)
  edge_gate (True)                 skip the edges none of whose pairs could act
                                   (an exact bound on the type-8 condition;
                                   triangle trees); False walks and tests every
                                   edge
  status()                         device status word (overflow flags)
  capture(n_substeps, n_iters)     one frame captured into a HIP graph for replay
  static_rebuild_every_frame (True)  rebuild the static tree at every forward as
                                   the reference does (False: build it once and
                                   again only when the padding changes)
  exclude_own_faces (True)         triangle trees: the cloth vertices' lists of
                                   the cloth tree leave out each vertex's own
                                   faces (which the narrow phase skips,
                                   XPBD.py:258) but count them toward the cap;
                                   positions and velocities are unchanged, the
                                   lists shorter (their counts then count the
                                   stored candidates)
  check_status (False)             read the status word after every forward
                                   and raise SlxError naming the flags set
  spring_launch (None)             (threads per block, vertices per thread) of
                                   the spring kernels, this solver only (None =
                                   the library default); walk_stack,
                                   active_grid, refit_shape likewise for the
                                   broad phase (results never depend on them)
Each Solver owns its side streams and events (slx_frame_ctx), so solvers in
different threads never share them.
`step` is an alias of `forward`.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from . import topology as _topo
from .broadphase import DeviceTree


def _f32(v) -> float:
    return float(np.float32(v))


class Solver:
    def __init__(self, mesh_dy, mesh_st, dHat, stiffness_stretch, stiffness_bending, g, dt):
        import torch
        if not torch.cuda.is_available():
            raise _abi.SlxError("Solver needs a HIP device (torch.cuda.is_available() is False)")
        _abi.lib()
        self.mesh_dy = mesh_dy
        self.mesh_st = mesh_st
        self.g = tuple(float(c) for c in g)
        self.dt = dt
        self.dHat = dHat
        self.stiffness_stretch = stiffness_stretch
        self.stiffness_bending = stiffness_bending
        self.damping = 0.001
        self.mu = 0.8
        self.padding = 0.05
        self.enable_velocity_update = False
        self.enable_collision_handling = False
        self.enable_move_obstacle = False
        self.export_mesh = False
        self.max_num_verts_dy = mesh_dy.num_verts
        self.max_num_edges_dy = mesh_dy.num_edges
        self.max_num_faces_dy = mesh_dy.num_faces
        self.max_num_verts_st = mesh_st.num_verts if mesh_st is not None else 0
        self.max_num_edges_st = mesh_st.num_edges if mesh_st is not None else 0
        self.max_num_faces_st = mesh_st.num_faces if mesh_st is not None else 0
        # extensions
        self.n_iters = 1
        self.bvh_mode = "cell"
        self.refit_every_substep = False
        self.cache_size = 40
        self.pair_cache_size = None
        self.candidate_cache = True
        self.cache_margin = None
        self.cache_capacity = None
        self.cache_every = 1
        self.static_cache_every = 4      # the static set's cache interval (None: cache_every)
        self.static_cache_margin = None  # and its margin (None: 2 x padding)
        self.enable_sewing = False
        self.sewing_stiffness = None
        self._sew = None
        self.enable_edge_edge = False
        self.ee_pair_cache_size = None
        self.edge_gate = True
        self.compliance_collision = 1e8          # XPBD.py:367
        self.records_per_vertex = 32
        self.method = "jacobi"                   # or "gs": graph-coloured Gauss-Seidel
        self.max_colors = 64
        self._gs = None
        dev = mesh_dy.x4.device
        self.device = dev
        nv = mesh_dy.num_verts
        self._y_alt = torch.empty((nv, 4), dtype=torch.float32, device=dev)
        self._status = torch.zeros(1, dtype=torch.int32, device=dev)
        self._cd = None            # collision buffers, allocated on first use
        self._ws = None
        self.check_status = False
        self.static_rebuild_every_frame = True
        self.spring_launch = None
        self.walk_stack = None
        self.active_grid = None
        self.refit_shape = None
        self.walk_wave = None
        self.exclude_own_faces = True
        self._ctx = None           # side streams + events of this solver

    # ------------------------------------------------------------ reference API
    def reset(self):
        self.mesh_dy.reset()
        if self.mesh_st is not None:
            self.mesh_st.reset()

    def set_fixed_vertices(self, fixed_vertices):
        """XPBD.py:389-395: fixed = 0 where selection >= 1, else 1."""
        import torch
        m = self.mesh_dy
        sel = torch.as_tensor(fixed_vertices, device=m.x4.device)
        m.fixed.copy_(torch.where(sel >= 1, 0.0, 1.0).to(torch.float32))
        _abi.call("slx_set_fixed", m.num_verts, _abi.ptr(m.fixed), _abi.ptr(m.m_inv), _abi.ptr(m.x4),
                  _abi.ptr(m.v4), _abi.ptr(m.y4), _abi.stream_handle())

    def scalars(self, n_substeps):
        """Per-substep scalars as the reference forms them: dt_sub and
        stiffness*dt_sub^2 in double, rounded to float32 at the kernel call
        (XPBD.py:446, 354-355 — the stiffness names are swapped there)."""
        dt_sub = self.dt / n_substeps
        return dict(dt=_f32(dt_sub), k_s=_f32(self.stiffness_bending * dt_sub * dt_sub),
                    k_b=_f32(self.stiffness_stretch * dt_sub * dt_sub), damping=_f32(self.damping),
                    g=tuple(_f32(c) for c in self.g))

    def forward(self, n_substeps, n_iters=None):
        n_iters = self.n_iters if n_iters is None else n_iters
        d = self._desc(n_substeps, n_iters)
        _abi.call("slx_frame", _abi.C.byref(d), _abi.stream_handle())
        if self.check_status:
            st = self.status()
            if st:
                raise _abi.SlxError(f"slx_frame status {st:#x}: {_abi.status_flags(st)}")

    def capture(self, n_substeps, n_iters=None):
        """One frame, forward(n_substeps, n_iters), captured into a HIP graph
        (torch.cuda.CUDAGraph) and returned: each graph.replay() runs that frame
        on the solver's current state with its launch shapes and buffers baked
        in, and no host work per launch.  An eager frame runs first (allocations,
        the static tree), so the state advances by one frame; the captured frame
        itself does not run.  Settings changed afterwards need a new capture;
        check_status is not consulted inside the graph."""
        import torch
        self.forward(n_substeps, n_iters)
        torch.cuda.synchronize()
        check, self.check_status = self.check_status, False
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        try:
            with torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    self.forward(n_substeps, n_iters)
        finally:
            self.check_status = check
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        return g

    def __del__(self):
        ctx, self._ctx = getattr(self, "_ctx", None), None
        if ctx:
            try:
                _abi.lib().slx_frame_ctx_destroy(ctx)
            except Exception:
                pass

    def _frame_ctx(self):
        if self._ctx is None:
            h = _abi.C.c_void_p()
            _abi.call("slx_frame_ctx_create", _abi.C.byref(h))
            self._ctx = h.value
        return self._ctx

    step = forward

    def load_sewing_pairs(self, path=None, pairs=None):
        """XPBD.py:405-417: pairs from animation/sewing.csv (default path) or an
        (n, 2) array."""
        import torch
        from . import selection, topology
        if pairs is None:
            pairs = selection.load_sewing_pairs(path or "animation/sewing.csv")
        pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        nv = self.mesh_dy.num_verts
        if len(pr) > nv:
            raise ValueError("more sewing pairs than dynamic vertices (the reference's field holds V)")
        verts, offs, ent = topology.sewing_lists(pr, nv)
        dev = self.device
        self._sew = dict(pairs=torch.from_numpy(pr.copy()).to(dev), verts=torch.from_numpy(verts).to(dev),
                         off=torch.from_numpy(offs).to(dev), list=torch.from_numpy(ent).to(dev),
                         ws=torch.zeros((max(len(verts), 1), 4), dtype=torch.float32, device=dev), n=len(verts),
                         np=len(pr))
        return pr

    def status(self) -> int:
        """Device status word (SLX_STATUS_* bits); synchronises."""
        return int(self._status.item())

    def clear_status(self):
        self._status.zero_()

    # --------------------------------------------------------------- internals
    def _collision_buffers(self):
        import torch
        pcap = self.cache_size if self.pair_cache_size is None else int(self.pair_cache_size)
        key = (self.bvh_mode, self.cache_size, self.records_per_vertex)
        if self._cd is not None and self._cd["key"] == key:
            self._pair_buffers(self._cd, pcap)
            self._cand_caches(self._cd)
            return self._cd
        m, s = self.mesh_dy, self.mesh_st
        nv, nvs = m.num_verts, (s.num_verts if s is not None else 0)
        dev = self.device
        cap = self.cache_size
        i32 = dict(dtype=torch.int32, device=dev)
        cd = dict(key=key)
        cd["dyn"] = DeviceTree(m, self.bvh_mode)
        cd["st"] = DeviceTree(s, self.bvh_mode) if s is not None else None
        cd["cand"] = torch.full((cap * (2 * nv + nvs) + 1,), -1, **i32)
        cd["cand_n"] = torch.zeros(2 * nv + nvs + 1, **i32)
        cd["pair_num"] = torch.zeros(2 * nv + nvs + 1, **i32)
        cd["rec_cap"] = int(self.records_per_vertex)
        cd["static_built"] = None          # padding the static tree was built with
        if s is not None:
            s.xs4 = s.x4
        self._cd = cd
        self._pair_buffers(cd, pcap)
        self._cand_caches(cd)
        return cd

    def _cand_caches(self, cd):
        """Candidate caches of the three query sets (triangle trees; batched trees
        walk from each query's instance root)."""
        from .broadphase import CandidateCache
        tri = self.bvh_mode == "tri"
        on = self.candidate_cache and tri
        ccap = (2 * self.cache_size if self.cache_capacity is None else int(self.cache_capacity)) if on else 0
        margin = _f32(0.25 * self.padding if self.cache_margin is None else self.cache_margin)
        smargin = _f32(2.0 * self.padding if self.static_cache_margin is None else self.static_cache_margin)
        key = (tri, ccap, margin, smargin)
        if cd.get("cache_key") == key:
            return
        cd["cache_key"] = key
        if not tri:                    # (the two-tier walk and the caches serve triangle trees)
            cd["caches"] = None
            return
        m, s = self.mesh_dy, self.mesh_st
        nv, nvs = m.num_verts, (s.num_verts if s is not None else 0)
        if on:
            cd["dyn"].enable_snapshot()
        cd["caches"] = [CandidateCache(nv, ccap, smargin, self.device) if s is not None else None,
                        CandidateCache(nv, ccap, margin, self.device),
                        CandidateCache(nvs, ccap, margin, self.device) if s is not None else None]

    def _edge_buffers(self, cd):
        """Edge-edge candidate lists, counts and pair records (allocated once the
        constraints are enabled)."""
        import torch
        m, s = self.mesh_dy, self.mesh_st
        ne, cap = m.num_edges, self.cache_size
        pcap = cap if self.ee_pair_cache_size is None else int(self.ee_pair_cache_size)
        key = (cap, pcap, self.enable_velocity_update, self.bvh_mode)
        if cd.get("ee_key") == key:
            return cd["ee"]
        i32 = dict(dtype=torch.int32, device=self.device)
        e = dict(ne=ne, cap=cap, pcap=pcap, fe_dy=m.face_edge_indices(),
                 fe_st=s.face_edge_indices() if s is not None else None,
                 c_st=torch.full((max(cap * ne, 1),), -1, **i32) if s is not None else None,
                 n_st=torch.zeros(max(ne, 1), **i32) if s is not None else None,
                 c_dy=torch.full((max(cap * ne, 1),), -1, **i32), n_dy=torch.zeros(max(ne, 1), **i32),
                 pair_num=torch.zeros(max(2 * ne, 1), **i32), order=m.edge_order(),
                 redo=torch.zeros(2 * (1 + ne), **i32) if self.bvh_mode == "tri" else None,
                 gate=torch.zeros(8, **i32), live=torch.zeros(2 * (1 + ne), **i32),
                 nes=s.num_edges if s is not None else 0)
        if self.enable_velocity_update:
            e["pair_cnt"] = torch.zeros(max(2 * ne, 1), **i32)
            e["pair_rec"] = torch.zeros(max(16 * pcap * 2 * ne, 16), dtype=torch.float32, device=self.device)
        cd["ee"], cd["ee_key"] = e, key
        return e

    def edge_pairs(self):
        """Last edge-edge pair counts (uncapped) and, with the velocity pass, the
        recorded pairs: dict with num [2ne] (static set, then dynamic set per
        edge), cnt [2ne], g [2ne, pcap, 4, 3], schur, edge (the other edge's id),
        type."""
        import torch
        e = self._cd["ee"]
        ne, pc = e["ne"], e["pcap"]
        out = dict(num=e["pair_num"][:2 * ne])
        if "pair_cnt" in e:
            r = e["pair_rec"][:16 * pc * 2 * ne].view(pc, 4, 2 * ne, 4).permute(2, 0, 1, 3)
            out.update(cnt=e["pair_cnt"][:2 * ne], g=r[..., :3], schur=r[:, :, 0, 3],
                       edge=r[:, :, 1, 3].contiguous().view(torch.int32),
                       type=r[:, :, 2, 3].contiguous().view(torch.int32))
        return out

    def _pair_buffers(self, cd, pcap):
        """Pair records for the velocity pass (allocated once it is enabled)."""
        import torch
        if not self.enable_velocity_update or cd.get("pair_cap") == pcap:
            return
        m, s = self.mesh_dy, self.mesh_st
        nq = 2 * m.num_verts + (s.num_verts if s is not None else 0)
        cd["pair_cap"] = pcap
        cd["pair_cnt"] = torch.zeros(nq + 1, dtype=torch.int32, device=self.device)
        cd["pair_rec"] = torch.zeros(max(16 * pcap * nq, 16), dtype=torch.float32, device=self.device)

    def _desc(self, n_substeps, n_iters):
        import torch
        m = self.mesh_dy
        sc = self.scalars(n_substeps)
        P = _abi.ptr
        d = _abi.FrameDesc()
        d.nv = m.num_verts
        d.x4, d.v4, d.y4, d.y4_alt = P(m.x4), P(m.v4), P(m.y4), P(self._y_alt)
        d.fixed, d.m_inv, d.faces = P(m.fixed), P(m.m_inv), P(m.face_indices)
        d.adj_slice_off, d.adj_slice_width, d.adj_entries = P(m.adj_slice_off), P(m.adj_slice_width), P(m.adj_entries)
        d.dt, d.k_s, d.k_b, d.damping = sc["dt"], sc["k_s"], sc["k_b"], sc["damping"]
        d.gx, d.gy, d.gz = sc["g"]
        d.padding, d.dhat, d.k_col = _f32(self.padding), _f32(self.dHat), _f32(self.compliance_collision)
        d.n_substeps, d.n_iters = n_substeps, n_iters
        d.collision = 1 if self.enable_collision_handling else 0
        d.refit_every_substep = 1 if self.refit_every_substep else 0
        d.status = P(self._status)
        d.ctx = self._frame_ctx()
        if self.spring_launch is not None:
            d.spring_block, d.spring_vpt = (int(v) for v in self.spring_launch)
        if self.walk_stack is not None:
            d.walk_stack = int(self.walk_stack)
        if self.active_grid is not None:
            d.active_grid = int(self.active_grid)
        if self.refit_shape is not None:
            d.refit_sub, d.refit_top = (int(v) for v in self.refit_shape)
        if self.walk_wave is not None:
            d.walk_wave = 1 if self.walk_wave else -1
        if self.method not in ("jacobi", "gs"):
            raise ValueError("method must be 'jacobi' or 'gs'")
        if self.method == "gs":
            gs = self.coloring()
            d.method, d.n_colors = 1, gs["n_colors"]
            d.gs_edges = P(gs["edges"])
            d.gs_color_off = gs["offsets"].ctypes.data_as(_abi.C.c_void_p)
        if self.enable_collision_handling:
            cd = self._collision_buffers()
            s = self.mesh_st
            d.nvs = s.num_verts if s is not None else 0
            if s is not None:
                d.xs4, d.faces_st = P(s.x4), P(s.face_indices)
                d.st = cd["st"].struct()
                # the reference rebuilds the static tree every frame with the current
                # padding (XPBD.py:454-458); with static_rebuild_every_frame False it
                # is built once and rebuilt only when the padding changes
                pad = _f32(self.padding)
                rebuild = self.static_rebuild_every_frame or cd["static_built"] != pad
                d.rebuild_static = 1 if rebuild else 0
                cd["static_built"] = pad
            d.dyn = cd["dyn"].struct()
            d.cap, d.rec_cap = self.cache_size, cd["rec_cap"]
            d.cand, d.cand_n, d.pair_num = P(cd["cand"]), P(cd["cand_n"]), P(cd["pair_num"])
            if self.exclude_own_faces and self.bvh_mode == "tri":
                if cd.get("own") is None:
                    cd["own"] = torch.from_numpy(_topo.own_faces(m.num_verts, m.fid_np)).to(self.device)
                d.own_faces = P(cd["own"])
            if cd.get("caches"):
                for k, c in enumerate(cd["caches"]):
                    if c is not None:
                        d.cache[k] = c.struct()
                d.cache_every = int(self.cache_every)
                d.cache_every_static = int(self.static_cache_every or 0)
            if self.enable_velocity_update:
                d.pair_cap, d.pair_cnt, d.pair_rec = cd["pair_cap"], P(cd["pair_cnt"]), P(cd["pair_rec"])
                if s is not None:
                    d.vs4, d.m_inv_st = P(s.v4), P(s.m_inv)
            if self.enable_edge_edge and m.num_edges > 0:
                if cd["dyn"].n_inst > 1:
                    raise ValueError("edge-edge constraints need a single (not batched) dynamic mesh")
                e = self._edge_buffers(cd)
                d.ee = _abi.EdgeSets(e["ne"], P(m.edge_indices), P(s.edge_indices) if s is not None else None,
                                     P(e["fe_dy"]), P(e["fe_st"]), e["cap"], P(e["c_st"]), P(e["n_st"]),
                                     P(e["c_dy"]), P(e["n_dy"]), e["pcap"], P(e["pair_num"]), P(e.get("pair_cnt")),
                                     P(e.get("pair_rec")), P(e["order"]), P(e["redo"]),
                                     P(e["gate"]) if self.edge_gate else None,
                                     P(e["live"]) if self.edge_gate else None, e["nes"])
        if self.enable_sewing and self._sew is not None and self._sew["np"] > 0:
            sw = self._sew
            dt_sub = self.dt / n_substeps
            ks = 0.5 * self.stiffness_bending if self.sewing_stiffness is None else self.sewing_stiffness
            d.n_sewn, d.sew_verts, d.sew_off, d.sew_list = sw["n"], P(sw["verts"]), P(sw["off"]), P(sw["list"])
            d.sew_pairs, d.sew_ws, d.k_sew = P(sw["pairs"]), P(sw["ws"]), _f32(ks * dt_sub * dt_sub)
        # velocity pass (XPBD.py:467-468)
        d.velocity = 1 if self.enable_velocity_update else 0
        d.mu = _f32(self.mu)
        need = _abi.lib().slx_frame_ws_bytes(_abi.C.byref(d))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        d.ws, d.ws_bytes = P(self._ws), self._ws.numel()
        self._last_desc = d
        return d

    def coloring(self):
        """Greedy stretch-edge colouring (graph_coloring.py:131-201), uploaded as
        int4 (v0, v1, l0 bits, 0) per edge in colour order."""
        import torch
        if self._gs is None:
            m = self.mesh_dy
            color, order, offsets = _abi.greedy_coloring(m.num_verts, m.eid_np, self.max_colors)
            e = m.eid_np[order]
            packed = np.zeros((len(order), 4), np.int32)
            packed[:, 0], packed[:, 1] = e[:, 0], e[:, 1]
            packed[:, 2] = m.l0_np[order].view(np.int32)
            self._gs = dict(color=color, order=order, offsets=np.ascontiguousarray(offsets, np.int32),
                            n_colors=len(offsets) - 1, edges=torch.from_numpy(packed).to(self.device))
        return self._gs

    # ----------------------------------------------------------- stage access
    def compute_y(self, dt):
        m = self.mesh_dy
        g = tuple(_f32(c) for c in self.g)
        _abi.call("slx_compute_y", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(m.y4),
                  _f32(dt), *g, _abi.stream_handle())

    def solve_spring_constraints(self, k_s, k_b):
        """One spring pass on y (through the alternate buffer)."""
        m = self.mesh_dy
        _abi.call("slx_spring_pass", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.fixed), _abi.ptr(m.y4),
                  _abi.ptr(self._y_alt), _abi.ptr(m.adj_slice_off), _abi.ptr(m.adj_slice_width),
                  _abi.ptr(m.adj_entries), _f32(k_s), _f32(k_b), _abi.stream_handle())
        m.y4.copy_(self._y_alt)

    def update_vx(self, dt):
        m = self.mesh_dy
        _abi.call("slx_update_vx", m.num_verts, _abi.ptr(m.x4), _abi.ptr(m.v4), _abi.ptr(m.y4),
                  _abi.ptr(m.fixed), _f32(self.damping), _f32(dt), _abi.stream_handle())

    def pairs(self):
        """Last recorded contacting pairs (velocity pass on): dict with cnt [nq],
        and per pair j < cnt[q] of query q: g [nq, pair_cap, 4, 3], schur, face,
        type (queries in the order vt_st, vt_dy, tv_st)."""
        import torch
        cd = self._cd
        pc = cd["pair_cap"]
        nq = cd["pair_cnt"].numel() - 1
        r = cd["pair_rec"][:16 * pc * nq].view(pc, 4, nq, 4).permute(2, 0, 1, 3)
        g = r[..., :3]
        return dict(cnt=cd["pair_cnt"][:nq], g=g, schur=r[:, :, 0, 3],
                    face=r[:, :, 1, 3].contiguous().view(torch.int32),
                    type=r[:, :, 2, 3].contiguous().view(torch.int32))

    def candidates(self):
        """Last broad-phase candidate lists: dict set -> (cand (cap, nq), counts)."""
        cd = self._cd
        m, s = self.mesh_dy, self.mesh_st
        nv, nvs, cap = m.num_verts, (s.num_verts if s is not None else 0), self.cache_size
        c, n = cd["cand"], cd["cand_n"]
        out = {"vt_dy": (c[cap * nv:2 * cap * nv].view(cap, nv), n[nv:2 * nv])}
        if nvs:
            out["vt_st"] = (c[:cap * nv].view(cap, nv), n[:nv])
            out["tv_st"] = (c[2 * cap * nv:2 * cap * nv + cap * nvs].view(cap, nvs), n[2 * nv:2 * nv + nvs])
        return out
