"""Algorithmic HBM byte model of the XPBD cloth step (FP32 / int32), the single
source of truth for roofline figures.

Each array is counted once per pass; on-chip temporaries (dx, nc) are not
counted.  V/E/B/F are the dynamic mesh's vertex, edge, bending and face counts,
Vs/Fs the static mesh's; cbar[set] is the measured mean number of broad-phase
candidates per query of a query set.

  predict                      40 V   (read x, v, f; write y)
  one Jacobi spring pass       12 (E+B) + 40 V   (ids + rest length; read y, x, f*w; write y)
  velocity / position update   52 V   (read x, y, f; write v, x)
  refit (face boxes fused)     160 F + 16 V  (faces + leaf order 16 F, topology 16 F, x 16 V; write 2F-1
                                              nodes 64 F and the 4-wide lines of the even-depth
                                              internal nodes, about half of them, 64 F; a batched
                                              tree writes every line: 224 F + 16 V)
  traversal (nq queries, tree of Ft faces)   16 nq + 4 nq + 4 c nq + tree, tree = 128 (Ft - 1) + 32 for the
                               4-wide node array, 32 (2 Ft - 1) for the binary node arrays
  narrow phase                 16 c V + 24 V  (per query set of V queries)
  contact apply                56 V   (counts, first-set sums, f; read/write y)
  per-frame rebuild            Morton 20 F + 12 V, radix sort 68 F, Karras 16 F

Candidate cache (triangle trees; cs[set] = mean leaves stored per query, the
cache's own count):
  cache build (a walk with the enlarged box, the exact list emitted on the way)
                               traversal(nq, Ft, c) + (16 + 4 + 4 cs) nq  (query snapshot,
                               count and leaves written)
  cache filter                 (16 + 16 + 4 + 4) nq + 4 cs nq + 32 Ft + (4 c + 4) nq
                               (query, snapshot, count, order; stored leaves; the leaf
                                boxes once; list and count written)

Velocity pass (enable_velocity_update), with P contacting pairs recorded by the
last collision pass of a substep and R bucket deltas (two per moving point of
each second- and third-set pair whose Cv < 0; the benchmark counts every such
pair, an upper bound):
  pair records (contacts)      64 P + 4 nq   (g0..g3, schur, face, type; counts)
  compute_velocity             52 V   (read x, y, f; write v)
  velocity contacts            4 nq + 76 P + 16 (V + Vs) + 8 V + 16 V + 20 R
                               (counts, records + face ids, velocities, f and w,
                                first-set sums, bucket deltas + counters)
  velocity apply               92 V + 16 R   (counts, sums, f, w; read/write v and x; buckets)

Edge-edge constraints (enable_edge_edge), ne dynamic edges, ce[set] mean
candidate faces per edge, Q edge pairs in contact (the position pass's
count), Re bucket deltas of the edge sets:
  edge traversal (per tree)    44 ne + 4 ce ne + tree   (edge ids 8, end points 32, count 4;
                               list; the tree as for the vertex queries)
  edge-edge contacts (per set) (48 + 40 ce) ne + 20 Re  (count, own ids and end points, pair
                               count; per candidate face its id, its sides' ids and their
                               end-point ids; bucket deltas + counters)
  bucket fold                  40 V + 16 Rv   (counts, running sums read and written; the
                               vertex-triangle records, Rv of them, read)
"""
from __future__ import annotations

HBM_PEAK_BYTES_PER_S = 8.0e12      # MI355X HBM3E


def predict(V):
    return 40 * V


def spring_pass(V, E, B):
    return 12 * (E + B) + 40 * V


def update_vx(V):
    return 52 * V


def refit(V, F, every_line=False):
    return (224 if every_line else 160) * F + 16 * V


def contact_apply(V):
    return 56 * V


def traversal(nq, Ft, c, quad=True):
    tree = 128 * max(Ft - 1, 0) + 32 if quad else 32 * (2 * Ft - 1)
    return 16 * nq + 4 * nq + 4 * c * nq + tree


def narrow(nq, c):
    return 16 * c * nq + 24 * nq


def rebuild(V, F):
    return (20 * F + 12 * V) + 68 * F + 16 * F


def cache_build(nq, Ft, c, cs, quad=True):
    return traversal(nq, Ft, c, quad) + (16 + 4 + 4 * cs) * nq


def cache_filter(nq, Ft, c, cs):
    return 40 * nq + 4 * cs * nq + 32 * Ft + (4 * c + 4) * nq


def pair_records(P, nq):
    return 64 * P + 4 * nq


def compute_velocity(V):
    return 52 * V


def vel_contacts(nq, P, V, Vs, R):
    return 4 * nq + 76 * P + 16 * (V + Vs) + 8 * V + 16 * V + 20 * R


def vel_apply(V, R):
    return 92 * V + 16 * R


def edge_traversal(ne, Ft, ce, quad=True):
    tree = 128 * max(Ft - 1, 0) + 32 if quad else 32 * (2 * Ft - 1)
    return 44 * ne + 4 * ce * ne + tree


def ee_contacts(ne, ce, Re=0):
    return (48 + 40 * ce) * ne + 20 * Re


def bucket_fold(V, Rv=0):
    return 40 * V + 16 * Rv


def kernel_bytes(name: str, n: dict):
    """Algorithmic bytes of one launch of kernel `name` (None if not modelled)."""
    V, E, B, F = n["V"], n["E"], n["B"], n["F"]
    c = n.get("cbar", {})
    sp = spring_pass(V, E, B)
    table = {
        "compute_y": predict(V),
        "spring": sp,
        "spring_pred": predict(V) + sp,
        "spring_final": sp + update_vx(V),
        "spring_pred_final": predict(V) + sp + update_vx(V),
        "update_vx": update_vx(V),
        "tri_refit": refit(V, F),
    }
    cs = n.get("cache", {})            # mean stored leaves per query of the cached sets
    if "Vs" in n:
        # the static (terrain) tree is coarse next to the queries: the wavefront walk over binary nodes
        table["traverse_vt_st"] = traversal(V, n["Fs"], c.get("vt_st", 0.0), quad=False)
        table["traverse_tv_st"] = traversal(n["Vs"], F, c.get("tv_st", 0.0))
        if "vt_st" in cs:
            table["traverse_vt_st"] = cache_filter(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"])
            table["cache_build_wave"] = cache_build(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"], quad=False)
    table["traverse_vt_dy"] = traversal(V, F, c.get("vt_dy", 0.0))
    if "vt_dy" in cs:
        table["traverse_vt_dy"] = cache_filter(V, F, c.get("vt_dy", 0.0), cs["vt_dy"])
        table["cache_build"] = cache_build(V, F, c.get("vt_dy", 0.0), cs["vt_dy"])
    nar = narrow(V, c.get("vt_dy", 0.0))
    if "Vs" in n:
        nar += narrow(V, c.get("vt_st", 0.0)) + narrow(n["Vs"], c.get("tv_st", 0.0))
    table["contacts"] = nar
    table["contact_apply"] = contact_apply(V)
    if "pairs" in n:
        Vs = n.get("Vs", 0)
        nq = 2 * V + Vs
        table["compute_velocity"] = compute_velocity(V)
        table["vel_contacts"] = vel_contacts(nq, n["pairs"], V, Vs, n["vel_records"])
        table["vel_apply"] = vel_apply(V, n["vel_records"])
    if "ee_cbar" in n:
        ce = n["ee_cbar"]
        table["traverse_ee_dy"] = edge_traversal(E, F, ce.get("ee_dy", 0.0))
        table["ee_contacts"] = sum(ee_contacts(E, ce[k]) for k in ("ee_st", "ee_dy") if k in ce)
        table["bucket_fold"] = bucket_fold(V)
        if "Fs" in n:
            table["traverse_ee_st"] = edge_traversal(E, n["Fs"], ce.get("ee_st", 0.0), quad=False)
    return table.get(name)


def step_bytes(n: dict, n_iters: int, substeps: int, collision: bool) -> float:
    """Bytes of one frame (`substeps` substeps)."""
    V, E, B, F = n["V"], n["E"], n["B"], n["F"]
    sub = predict(V) + n_iters * spring_pass(V, E, B) + update_vx(V)
    total = substeps * sub
    if collision:
        c = n.get("cbar", {})
        cs = n.get("cache", {})
        per_iter = narrow(V, c.get("vt_dy", 0.0)) + contact_apply(V)
        walks = n_iters * traversal(V, F, c.get("vt_dy", 0.0))
        if "vt_dy" in cs:              # one build and (n_iters - 1) filters per substep
            walks = (cache_build(V, F, c.get("vt_dy", 0.0), cs["vt_dy"]) +
                     (n_iters - 1) * cache_filter(V, F, c.get("vt_dy", 0.0), cs["vt_dy"]))
        if "Vs" in n:
            per_iter += narrow(V, c.get("vt_st", 0.0)) + narrow(n["Vs"], c.get("tv_st", 0.0))
            if "vt_st" in cs:
                walks += (cache_build(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"], quad=False) +
                          (n_iters - 1) * cache_filter(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"]))
            else:
                walks += n_iters * traversal(V, n["Fs"], c.get("vt_st", 0.0), quad=False)
            total += substeps * traversal(n["Vs"], F, c.get("tv_st", 0.0))
        total += substeps * (refit(V, F) + n_iters * per_iter + walks) + rebuild(V, F)
    if "pairs" in n:
        # velocity pass: compute_velocity + contacts + apply (update_x fused) replace update_vx
        Vs = n.get("Vs", 0)
        nq = 2 * V + Vs
        R = n["vel_records"]
        total += substeps * (pair_records(n["pairs"], nq) + compute_velocity(V) +
                             vel_contacts(nq, n["pairs"], V, Vs, R) + vel_apply(V, R) - update_vx(V))
    if collision and "ee_cbar" in n:
        # edge-edge constraints: per pass the edge queries of both trees, the two sets, the fold
        ce = n["ee_cbar"]
        per = edge_traversal(E, F, ce.get("ee_dy", 0.0)) + ee_contacts(E, ce.get("ee_dy", 0.0)) + bucket_fold(V)
        if "Fs" in n:
            per += edge_traversal(E, n["Fs"], ce.get("ee_st", 0.0), quad=False) + ee_contacts(E, ce.get("ee_st", 0.0))
        total += substeps * n_iters * per
    return total


def substep_bytes(V, E, B, n_iters=1):
    return predict(V) + n_iters * spring_pass(V, E, B) + update_vx(V)
