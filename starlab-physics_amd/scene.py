"""Scene geometry: cloth grids, icospheres, height-field terrain, OBJ I/O and
mesh concatenation with per-mesh vertex offsets.

The reference loads OBJ assets through meshtaichi_patcher / libigl
(framework/meshio/meshTaichiWrapper.py:25, framework/meshio/concat.py:11-58);
the synthetic workloads below are generated instead, seeded with
numpy.random.default_rng(seed).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np


@dataclass
class TriMesh:
    x: np.ndarray          # (V,3) float32
    faces: np.ndarray      # (F,3) int32


def grid_cloth(n: int, size: float = 1.0, height: float = 0.5,
               center=(0.0, 0.0), jitter: float = 0.0, seed: int = 0) -> TriMesh:
    """n x n vertex cloth in the XZ plane at y = height, centred on
    (center[0], center[1]).  Vertex (i, j) has id i*n + j, x grows with i,
    z with j.  Quad (i, j) is split along (i,j)-(i+1,j+1)."""
    t = np.linspace(-0.5 * size, 0.5 * size, n, dtype=np.float64)
    xi, zj = np.meshgrid(t + center[0], t + center[1], indexing="ij")
    x = np.stack([xi, np.full_like(xi, height), zj], axis=-1).reshape(-1, 3)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter, jitter, size=x.shape)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).reshape(-1)
    b = ((i + 1) * n + j).reshape(-1)
    c = ((i + 1) * n + j + 1).reshape(-1)
    d = (i * n + j + 1).reshape(-1)
    f = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    return TriMesh(x.astype(np.float32), f.astype(np.int32))


def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> TriMesh:
    """Icosahedron refined `subdiv` times by edge midpoints, projected to the
    sphere.  subdiv=4 gives 2562 vertices / 5120 faces."""
    p = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0),
         (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    verts = [np.array(q, np.float64) / np.linalg.norm(q) for q in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    x = np.array(verts) * radius + np.asarray(center, np.float64)
    return TriMesh(x.astype(np.float32), np.array(faces, np.int32))


def terrain_height(x, z, waves: int = 8, amplitude: float = 0.1, base: float = 0.0, seed: int = 0):
    """y = base + sum_k (amplitude/waves) sin(kx_k x + kz_k z + phi_k) (float64);
    wave vectors have magnitude U(2, 12) rad/m and uniform direction; phases U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    mag = rng.uniform(2.0, 12.0, waves)
    ang = rng.uniform(0.0, 2 * np.pi, waves)
    phi = rng.uniform(0.0, 2 * np.pi, waves)
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    h = np.zeros(np.broadcast(x, z).shape)
    for k in range(waves):
        h += (amplitude / waves) * np.sin(mag[k] * np.cos(ang[k]) * x + mag[k] * np.sin(ang[k]) * z + phi[k])
    return base + h


def terrain(m: int = 142, size: float = 1.2, waves: int = 8, amplitude: float = 0.1,
            base: float = 0.0, seed: int = 0) -> TriMesh:
    """m x m height field over [-size/2, size/2]^2 of terrain_height."""
    g = grid_cloth(m, size=size, height=0.0)
    x = g.x.astype(np.float64)
    x[:, 1] = terrain_height(x[:, 0], x[:, 2], waves, amplitude, base, seed)
    return TriMesh(x.astype(np.float32), g.faces)


def stacked_layers(n: int, layers: int, spacing: float, size: float = 1.0,
                   height: float = 0.5, jitter: float = 0.0, seed: int = 0) -> TriMesh:
    """`layers` copies of an n x n cloth stacked `spacing` apart in y,
    concatenated with per-layer vertex offsets (framework/meshio/concat.py:11-58)."""
    parts = [grid_cloth(n, size=size, height=height + k * spacing) for k in range(layers)]
    m = concat(parts)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        m.x = (m.x.astype(np.float64) + rng.uniform(-jitter, jitter, m.x.shape)).astype(np.float32)
    return m


def concat(meshes) -> TriMesh:
    """Concatenate meshes; faces are offset by the running vertex count
    (framework/meshio/concat.py:17-48)."""
    xs, fs, off = [], [], 0
    for m in meshes:
        xs.append(m.x)
        fs.append(m.faces + off)
        off += len(m.x)
    return TriMesh(np.concatenate(xs).astype(np.float32), np.concatenate(fs).astype(np.int32))


def concat_offsets(meshes):
    off, out = 0, []
    for m in meshes:
        out.append(off)
        off += len(m.x)
    out.append(off)
    return out


def read_obj(path: str, dtype=np.float32) -> TriMesh:
    """Vertices and faces of an OBJ file (polygons are fan-triangulated,
    1-based and negative indices accepted, texture/normal indices ignored), the
    faces libigl's read_obj returns (framework/meshio/concat.py:20).  Positions
    are parsed as doubles and returned as `dtype`."""
    vs, fs = [], []
    with open(path, "r") as fh:
        for line in fh:
            if line.startswith("v "):
                vs.append([float(t) for t in line.split()[1:4]])
            elif line.startswith("f "):
                idx = []
                for tok in line.split()[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(vs) + k)
                for t in range(1, len(idx) - 1):
                    fs.append([idx[0], idx[t], idx[t + 1]])
    return TriMesh(np.array(vs, np.float64).astype(dtype).reshape(-1, 3), np.array(fs, np.int32).reshape(-1, 3))


def write_obj(path: str, x: np.ndarray, faces: np.ndarray) -> None:
    """Vertices and triangles as OBJ text (1-based faces).  Positions keep every
    bit: float32 arrays are written with 9 significant digits, float64 with 17,
    so reading the file back gives the same values."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    x = np.asarray(x)
    fmt = "v %.17g %.17g %.17g\n" if x.dtype == np.float64 else "v %.9g %.9g %.9g\n"
    with open(path, "w") as fh:
        for p in x.astype(np.float64):
            fh.write(fmt % (p[0], p[1], p[2]))
        for f in np.asarray(faces, np.int64):
            fh.write("f %d %d %d\n" % (f[0] + 1, f[1] + 1, f[2] + 1))


def concat_mesh(concat_model_name, model_dir, model_names, translations, scales, out_dir=None):
    """framework/meshio/concat.py:11-58 concat_mesh: read each model (doubles),
    scale it, translate it, offset its faces by the running vertex count, write
    the concatenation as OBJ `concat_model_name` and return the vertex offsets
    [0, n0, n0 + n1, ..., total].  The reference writes into its models/OBJ
    folder; here into out_dir (default model_dir)."""
    offsets, xs, fs, acc = [], [], [], 0
    for name, t, sc in zip(model_names, translations, scales):
        m = read_obj(os.path.join(model_dir, name), dtype=np.float64)
        offsets.append(acc)
        xs.append(sc * m.x + np.asarray(t, np.float64))
        fs.append(m.faces.astype(np.int64) + acc)
        acc += len(m.x)
    offsets.append(acc)
    x = np.concatenate(xs, axis=0) if xs else np.zeros((0, 3))
    f = np.concatenate(fs, axis=0) if fs else np.zeros((0, 3), np.int64)
    write_obj(os.path.join(model_dir if out_dir is None else out_dir, concat_model_name), x, f)
    return offsets


def rotation_matrix(rot_deg) -> np.ndarray:
    """R = Rz(rz) Ry(ry) Rx(rx) from Euler angles in degrees
    (the rotation applied by meshTaichiWrapper.applyTransform, :315-332)."""
    ax, ay, az = (math.radians(float(a)) for a in rot_deg)
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def apply_transform(x: np.ndarray, trans=(0, 0, 0), rot=(0, 0, 0), scale=1.0) -> np.ndarray:
    """scale, then rotate, then translate, in float32 (meshTaichiWrapper.py:315-332)."""
    x = np.asarray(x, np.float32) * np.float32(scale)
    if any(float(r) != 0.0 for r in rot):
        x = (x.astype(np.float64) @ rotation_matrix(rot).T).astype(np.float32)
    return (x + np.asarray(trans, np.float32)).astype(np.float32)
