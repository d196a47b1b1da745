"""Mesh: the reference's MeshTaichiWrapper surface backed by PyTorch-ROCm tensors.

Mirrors framework/meshio/meshTaichiWrapper.py:11-112 — constructor
(model_dir, model_name, offsets, trans, rot, scale, is_static), per-vertex fields
verts.{x, v, y, fixed, m_inv, x0}, edges.l0, face_indices, edge_indices,
bending_indices / bending_l0 / bending_constraint_count, reset(), computeAABB(),
computeAABB_faces(), export().

Device layout: x4/y4/v4 are (V,4) float32 tensors holding (position, fixed*m_inv)
and (velocity, fixed); verts.x / verts.v / verts.y are the (V,3) views.
"""
from __future__ import annotations

import os
import types

import numpy as np

from . import scene as _scene
from . import topology as _topo


class Mesh:
    def __init__(self, model_dir=None, model_name=None, offsets=(0,), trans=(0.0, 0.0, 0.0),
                 rot=(0.0, 0.0, 0.0), scale=1.0, is_static=False, *, x=None, faces=None,
                 device="cuda", topology=None):
        if x is None:
            if model_dir is None or model_name is None:
                raise ValueError("Mesh needs (model_dir, model_name) or x= and faces=")
            m = _scene.read_obj(os.path.join(model_dir, model_name))
            x, faces = m.x, m.faces
        self.is_static = is_static
        self.offsets = list(offsets)
        self.trans, self.rot, self.scale = trans, rot, scale
        x = _scene.apply_transform(np.asarray(x, np.float32), trans, rot, scale)
        self.fid_np = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
        self.x0_np = np.ascontiguousarray(x, np.float32)
        self.num_verts = len(self.x0_np)
        self.num_faces = len(self.fid_np)
        if topology is not None:
            # precomputed (edges, l0, m_inv, bending pairs, bending l0), e.g. per packed instance
            self.eid_np, self.l0_np, self.m_inv_np, self.bend_np, self.bending_l0_np = topology
        else:
            self.eid_np = _topo.edges_from_faces(self.fid_np)
            self.l0_np = _topo.rest_lengths(self.x0_np, self.eid_np)
            self.m_inv_np = _topo.lumped_inverse_mass(self.num_verts, self.eid_np, self.l0_np)
            if is_static:
                self.bend_np = np.zeros((0, 2), np.int32)
            else:
                self.bend_np = _topo.bending_pairs(self.fid_np)
            self.bending_l0_np = _topo.rest_lengths(self.x0_np, self.bend_np)
        self.num_edges = len(self.eid_np)
        self.bending_constraint_count = len(self.bend_np)
        self.adj = None if is_static else _topo.Adjacency(
            self.num_verts, self.eid_np, self.l0_np, self.bend_np, self.bending_l0_np)
        self.device = device
        self._upload(device)

    # ------------------------------------------------------------------ device
    def _upload(self, device):
        import torch
        self.torch_device = torch.device(device) if device is not None else None
        if device is None:
            return
        dev = self.torch_device
        nv = self.num_verts
        f32 = torch.float32
        self.x0 = torch.from_numpy(self.x0_np).to(dev)
        self.m_inv = torch.from_numpy(self.m_inv_np).to(dev)
        self.fixed = torch.ones(nv, dtype=f32, device=dev)
        self.x4 = torch.zeros((nv, 4), dtype=f32, device=dev)
        self.v4 = torch.zeros((nv, 4), dtype=f32, device=dev)
        self.y4 = torch.zeros((nv, 4), dtype=f32, device=dev)
        self.x4[:, :3] = self.x0
        self.y4[:, :3] = self.x0
        self._refresh_products()
        self.vert_order = torch.from_numpy(_topo.morton_order(self.x0_np)).to(dev)
        self.face_indices = torch.from_numpy(self.fid_np.reshape(-1)).to(dev)
        self.edge_indices = torch.from_numpy(self.eid_np.reshape(-1)).to(dev)
        self.bending_indices = torch.from_numpy(self.bend_np.reshape(-1)).to(dev)
        self.bending_l0 = torch.from_numpy(self.bending_l0_np).to(dev)
        self.edges = types.SimpleNamespace(l0=torch.from_numpy(self.l0_np).to(dev))
        self.faces = types.SimpleNamespace(aabb_min=None, aabb_max=None)
        if self.adj is not None:
            self.adj_slice_off = torch.from_numpy(self.adj.slice_off).to(dev)
            self.adj_slice_width = torch.from_numpy(self.adj.slice_width).to(dev)
            self.adj_entries = torch.from_numpy(self.adj.entries.reshape(-1)).to(dev)
        self.verts = types.SimpleNamespace(
            x=self.x4[:, :3], v=self.v4[:, :3], y=self.y4[:, :3], fixed=self.fixed,
            m_inv=self.m_inv, x0=self.x0)

    def face_edge_indices(self):
        """framework/meshio/meshTaichiWrapper.py:115-119 face_edge_indices: the
        edge ids of each face's sides (v0v1, v1v2, v2v0), [F*3] int32 on the
        device, built on first use."""
        fe = getattr(self, "_face_edges", None)
        if fe is None:
            import torch
            self.fe_np = _topo.face_edges(self.fid_np, self.eid_np)
            fe = torch.from_numpy(self.fe_np.reshape(-1)).to(self.torch_device)
            self._face_edges = fe
        return fe

    def edge_order(self):
        """Edges in a spatially coherent order (by the Morton rank of their first
        vertex), [E] int32 on the device, built on first use."""
        eo = getattr(self, "_edge_order", None)
        if eo is None:
            import torch
            vo = self.vert_order.cpu().numpy()
            eo = torch.from_numpy(_topo.edge_order(self.eid_np, vo)).to(self.torch_device)
            self._edge_order = eo
        return eo

    def _refresh_products(self):
        """x4.w = y4.w = fixed*m_inv, v4.w = fixed (the products the kernels read)."""
        fw = self.fixed * self.m_inv
        self.x4[:, 3] = fw
        self.y4[:, 3] = fw
        self.v4[:, 3] = self.fixed

    # ------------------------------------------------------------ reference API
    def reset(self):
        """meshTaichiWrapper.py:230-233: x = x0, v = 0, fixed = 0 (all pinned until
        set_fixed_vertices is called, as in the reference)."""
        self.x4[:, :3] = self.x0
        self.v4[:, :3] = 0.0
        self.fixed.fill_(0.0)
        self._refresh_products()

    def computeAABB(self, padding):
        """meshTaichiWrapper.py:334-347 (mesh box of x, minus/plus padding)."""
        x = self.x4[:, :3]
        lo = torch_min(x) - padding
        hi = torch_max(x) + padding
        return lo, hi

    def computeAABB_faces(self, padding):
        """meshTaichiWrapper.py:350-361 (per-face box of x, minus/plus padding)."""
        f = self.face_indices.view(-1, 3).long()
        x = self.x4[:, :3]
        p = x[f]
        self.faces.aabb_min = p.amin(dim=1) - padding
        self.faces.aabb_max = p.amax(dim=1) + padding
        return self.faces.aabb_min, self.faces.aabb_max

    def positions(self) -> np.ndarray:
        return self.x4[:, :3].detach().cpu().numpy()

    def export(self, scene_name, frame, is_static=False, root="../results"):
        """meshTaichiWrapper.py:369-386: write the current x as an OBJ."""
        sub = "StaticMesh_ID_" if is_static else "Mesh_ID_"
        path = os.path.join(root, scene_name, sub, "Mesh_obj_" + str(frame) + ".obj")
        _scene.write_obj(path, self.positions(), self.fid_np)
        return path

    def host_dict(self):
        """Topology and state as numpy arrays (the oracle's input format)."""
        x = self.positions() if self.device is not None else self.x0_np
        fixed = self.fixed.cpu().numpy() if self.device is not None else np.ones(self.num_verts, np.float32)
        return dict(x=x.astype(np.float32), x0=self.x0_np, faces=self.fid_np, edges=self.eid_np,
                    l0=self.l0_np, bend=self.bend_np, bl0=self.bending_l0_np, fixed=fixed,
                    m_inv=self.m_inv_np, fe=_topo.face_edges(self.fid_np, self.eid_np))

    @classmethod
    def from_trimesh(cls, tm, is_static=False, device="cuda", **kw):
        return cls(x=tm.x, faces=tm.faces, is_static=is_static, device=device, **kw)

    @classmethod
    def from_instances(cls, parts, is_static=False, device="cuda"):
        """Pack independent meshes into one, CSR-style: vertex and face offsets per
        instance (the concatenation of framework/meshio/concat.py:11-58), plus the
        instance id of every vertex and face.  Constraints never cross instances,
        so the solver runs on the packed mesh unchanged; the broad phase keeps one
        subtree per instance (slx_seg_build)."""
        tm = _scene.concat(parts)
        # topology per instance, offset and concatenated: instance vertex ranges are
        # increasing, so this equals the topology of the packed mesh, in the same order
        E, L, W, Bp, BL = [], [], [], [], []
        off = 0
        for p in parts:
            px = np.ascontiguousarray(p.x, np.float32)
            e = _topo.edges_from_faces(p.faces)
            l0 = _topo.rest_lengths(px, e)
            b = np.zeros((0, 2), np.int32) if is_static else _topo.bending_pairs(p.faces)
            E.append(e + off)
            L.append(l0)
            W.append(_topo.lumped_inverse_mass(len(px), e, l0))
            Bp.append(b + off)
            BL.append(_topo.rest_lengths(px, b))
            off += len(p.x)
        topo = (np.concatenate(E).astype(np.int32), np.concatenate(L), np.concatenate(W),
                np.concatenate(Bp).astype(np.int32).reshape(-1, 2), np.concatenate(BL))
        m = cls(x=tm.x, faces=tm.faces, is_static=is_static, device=device, topology=topo)
        nvs = np.array([len(p.x) for p in parts], np.int64)
        nfs = np.array([len(p.faces) for p in parts], np.int64)
        ids = np.arange(len(parts), dtype=np.int32)
        m.inst_np = dict(n=len(parts),
                         vert_base=np.concatenate([[0], np.cumsum(nvs)]).astype(np.int32),
                         face_base=np.concatenate([[0], np.cumsum(nfs)]).astype(np.int32),
                         face_inst=np.repeat(ids, nfs).astype(np.int32),
                         vert_inst=np.repeat(ids, nvs).astype(np.int32))
        if device is not None:
            import torch
            m.inst = {k: (torch.from_numpy(v).to(m.torch_device) if isinstance(v, np.ndarray) else v)
                      for k, v in m.inst_np.items()}
        return m

    def __len__(self):
        return self.num_verts


def torch_min(x):
    return x.amin(dim=0)


def torch_max(x):
    return x.amax(dim=0)
