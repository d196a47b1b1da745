"""HBM byte models of the XPBD cloth step (FP32 / int32): the single source of
truth for roofline figures.

Two models, kept apart:

ALGORITHMIC bytes (`kernel_bytes`, `step_bytes`) -- what the step must move
whatever the layout: each array counted once per pass, on-chip temporaries
(dx, nc) not counted.  These are the figures roofline fractions are quoted on.
V/E/B/F are the dynamic mesh's vertex, edge, bending and face counts, Vs/Fs the
static mesh's, c[set] the measured mean number of broad-phase candidates per
query of a query set (the reference's full lists, key "cbar"):

  predict (compute_y)          40 V           read x, v, f; write y
  one Jacobi iteration         12 (E+B) + 40 V  ids + rest length; read y, x, f*w; write y
  finalize (velocity, update_x) 52 V          read x, y, f; write v, x
  refit                        76 F + 12 V    faces 12 F; x 12 V; write 2F-1 node boxes (~64 F)
  traversal (nq queries, Ft)   16 nq + 4 c nq + 32 (2 Ft - 1)
                                              query boxes (a position + its padding), the
                                              list written, one pass over the node boxes
  narrow phase (per set)       16 c nq + 24 nq
  per-frame rebuild            Morton 20 F + 12 V, radix sort 68 F, Karras 16 F

and, where the table above has no row:

  update_dx after the collision pass   28 V   read y, f; write y
  velocity-level contact pass (pair records from the last collision pass, P
  contacting pairs, R bucket deltas):
    pair records               64 P + 4 nq
    velocity contacts          4 nq + 76 P + 16 (V + Vs) + 8 V + 16 V + 20 R
    velocity apply             92 V + 16 R
  edge-edge sets (ne edges, ce mean candidate faces, Re bucket deltas):
    edge traversal             44 ne + 4 ce ne + 32 (2 Ft - 1)
    edge-edge contacts         (48 + 40 ce) ne + 20 Re
    bucket fold                40 V + 16 Rv

The candidate cache is an implementation device: a walk that builds it and a
filter that reads it are both, algorithmically, one traversal of the query set.

IMPLEMENTATION bytes (`kernel_impl_bytes`, `step_impl_bytes`) -- what this
package's layout moves at best, with c the mean length of the lists as stored
(key "cbar_stored": the cloth vertices' lists of the cloth tree leave out their
own faces): float4 vertex records (16 B per position), the
4-wide node lines a walk loads (128 B per line, 128 (Ft - 1) + 32 for the
array), the even-depth 4-wide lines the refit writes besides the binary node
arrays, the cache's stored leaves and snapshots:

  predict 40 V; Jacobi 12 (E+B) + 40 V; finalize 52 V
  refit                        160 F + 16 V   (224 F + 16 V with every 4-wide line, batched trees)
  traversal                    20 nq + 4 c nq + 128 (Ft - 1) + 32 (4-wide) / 32 (2 Ft - 1) (binary)
  cache build                  traversal + (20 + 4 cs) nq   (snapshot, count, stored leaves)
  cache filter                 40 nq + 4 cs nq + 32 Ft + (4 c + 4) nq
  narrow phase                 16 c nq + 24 nq; contact apply 56 V (counts, first-set sums, f, y)
"""
from __future__ import annotations

HBM_PEAK_BYTES_PER_S = 8.0e12      # MI355X HBM3E


# ------------------------------------------------------------------ algorithmic

def predict(V):
    return 40 * V


def spring_pass(V, E, B):
    return 12 * (E + B) + 40 * V


def update_vx(V):
    return 52 * V


def refit(V, F):
    return 76 * F + 12 * V


def traversal(nq, Ft, c):
    return 16 * nq + 4 * c * nq + 32 * max(2 * Ft - 1, 0)


def narrow(nq, c):
    return 16 * c * nq + 24 * nq


def update_dx(V):
    return 28 * V


def rebuild(V, F):
    return (20 * F + 12 * V) + 68 * F + 16 * F


def pair_records(P, nq):
    return 64 * P + 4 * nq


def compute_velocity(V):
    return 52 * V


def vel_contacts(nq, P, V, Vs, R):
    return 4 * nq + 76 * P + 16 * (V + Vs) + 8 * V + 16 * V + 20 * R


def vel_apply(V, R):
    return 92 * V + 16 * R


def edge_traversal(ne, Ft, ce):
    return 44 * ne + 4 * ce * ne + 32 * max(2 * Ft - 1, 0)


def ee_contacts(ne, ce, Re=0):
    return (48 + 40 * ce) * ne + 20 * Re


def bucket_fold(V, Rv=0):
    return 40 * V + 16 * Rv


# --------------------------------------------------------------- implementation

def impl_refit(V, F, every_line=False):
    return (224 if every_line else 160) * F + 16 * V


def impl_traversal(nq, Ft, c, quad=True):
    tree = 128 * max(Ft - 1, 0) + 32 if quad else 32 * (2 * Ft - 1)
    return 16 * nq + 4 * nq + 4 * c * nq + tree


def impl_cache_build(nq, Ft, c, cs, quad=True):
    return impl_traversal(nq, Ft, c, quad) + (16 + 4 + 4 * cs) * nq


def impl_cache_filter(nq, Ft, c, cs):
    return 40 * nq + 4 * cs * nq + 32 * Ft + (4 * c + 4) * nq


def impl_contact_apply(V):
    return 56 * V


def impl_edge_traversal(ne, Ft, ce, quad=True):
    tree = 128 * max(Ft - 1, 0) + 32 if quad else 32 * (2 * Ft - 1)
    return 44 * ne + 4 * ce * ne + tree


# ------------------------------------------------------------------ per kernel

def _table(n: dict, impl: bool):
    V, E, B, F = n["V"], n["E"], n["B"], n["F"]
    c = n.get("cbar", {})
    if impl:                           # the lists as stored (without the cloth vertices' own faces)
        c = n.get("cbar_stored", c)
    cs = n.get("cache", {})            # mean stored leaves per query of the cached sets
    sp = spring_pass(V, E, B)
    t = {
        "compute_y": predict(V),
        "spring": sp,
        "spring_pred": predict(V) + sp,
        "spring_final": sp + update_vx(V),
        "spring_pred_final": predict(V) + sp + update_vx(V),
        "update_vx": update_vx(V),
        "tri_refit": impl_refit(V, F) if impl else refit(V, F),
    }

    def walk(nq, Ft, cc, quad=True):
        return impl_traversal(nq, Ft, cc, quad) if impl else traversal(nq, Ft, cc)

    if "Vs" in n:
        # the static (terrain) tree is coarse next to the queries: the wavefront walk over binary nodes
        t["traverse_vt_st"] = walk(V, n["Fs"], c.get("vt_st", 0.0), quad=False)
        t["traverse_tv_st"] = walk(n["Vs"], F, c.get("tv_st", 0.0))
        if "vt_st" in cs:
            t["cache_build_wave"] = (impl_cache_build(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"], quad=False)
                                     if impl else traversal(V, n["Fs"], c.get("vt_st", 0.0)))
            t["traverse_vt_st"] = (impl_cache_filter(V, n["Fs"], c.get("vt_st", 0.0), cs["vt_st"])
                                   if impl else traversal(V, n["Fs"], c.get("vt_st", 0.0)))
    t["traverse_vt_dy"] = walk(V, F, c.get("vt_dy", 0.0))
    if "vt_dy" in cs:
        t["cache_build"] = (impl_cache_build(V, F, c.get("vt_dy", 0.0), cs["vt_dy"])
                            if impl else traversal(V, F, c.get("vt_dy", 0.0)))
        t["traverse_vt_dy"] = (impl_cache_filter(V, F, c.get("vt_dy", 0.0), cs["vt_dy"])
                               if impl else traversal(V, F, c.get("vt_dy", 0.0)))
    t["contacts"] = narrow(V, c.get("vt_dy", 0.0))
    if "Vs" in n:
        # the first set's narrow phase runs as a launch of its own (contacts_vt_st)
        t["contacts_vt_st"] = narrow(V, c.get("vt_st", 0.0))
        t["contacts"] += narrow(n["Vs"], c.get("tv_st", 0.0))
    t["contact_apply"] = impl_contact_apply(V) if impl else update_dx(V)
    if "pairs" in n:
        Vs = n.get("Vs", 0)
        nq = 2 * V + Vs
        t["compute_velocity"] = compute_velocity(V)
        t["vel_contacts"] = vel_contacts(nq, n["pairs"], V, Vs, n["vel_records"])
        t["vel_apply"] = vel_apply(V, n["vel_records"])
    if "ee_cbar" in n:
        ce = n["ee_cbar"]
        et = (lambda ne, Ft, cc, quad=True: impl_edge_traversal(ne, Ft, cc, quad)) if impl else \
             (lambda ne, Ft, cc, quad=True: edge_traversal(ne, Ft, cc))
        t["traverse_ee_dy"] = et(E, F, ce.get("ee_dy", 0.0))
        t["ee_contacts"] = sum(ee_contacts(E, ce[k]) for k in ("ee_st", "ee_dy") if k in ce)
        t["bucket_fold"] = bucket_fold(V)
        if "Fs" in n:
            t["traverse_ee_st"] = et(E, n["Fs"], ce.get("ee_st", 0.0), False)
    return t


def kernel_bytes(name: str, n: dict):
    """Algorithmic bytes of one launch of kernel `name` (None if not modelled)."""
    return _table(n, False).get(name)


def kernel_impl_bytes(name: str, n: dict):
    """Implementation (layout) bytes of one launch of kernel `name`."""
    return _table(n, True).get(name)


# ------------------------------------------------------------------- per frame

def step_bytes(n: dict, n_iters: int, substeps: int, collision: bool) -> float:
    """Algorithmic bytes of one frame (`substeps` substeps): per substep predict,
    n_iters x (Jacobi pass + three query sets' traversals + narrow phases +
    update_dx), finalize, the refit; the rebuild once per frame."""
    V, E, B, F = n["V"], n["E"], n["B"], n["F"]
    total = substeps * (predict(V) + n_iters * spring_pass(V, E, B) + update_vx(V))
    if collision:
        c = n.get("cbar", {})
        per_iter = traversal(V, F, c.get("vt_dy", 0.0)) + narrow(V, c.get("vt_dy", 0.0)) + update_dx(V)
        if "Vs" in n:
            per_iter += (traversal(V, n["Fs"], c.get("vt_st", 0.0)) + narrow(V, c.get("vt_st", 0.0)) +
                         traversal(n["Vs"], F, c.get("tv_st", 0.0)) + narrow(n["Vs"], c.get("tv_st", 0.0)))
        total += substeps * (refit(V, F) + n_iters * per_iter) + rebuild(V, F)
    if "pairs" in n:
        # velocity pass: compute_velocity + contacts + apply (update_x fused) replace the finalize
        Vs = n.get("Vs", 0)
        nq = 2 * V + Vs
        R = n["vel_records"]
        total += substeps * (pair_records(n["pairs"], nq) + compute_velocity(V) +
                             vel_contacts(nq, n["pairs"], V, Vs, R) + vel_apply(V, R) - update_vx(V))
    if collision and "ee_cbar" in n:
        ce = n["ee_cbar"]
        per = edge_traversal(E, F, ce.get("ee_dy", 0.0)) + ee_contacts(E, ce.get("ee_dy", 0.0)) + bucket_fold(V)
        if "Fs" in n:
            per += edge_traversal(E, n["Fs"], ce.get("ee_st", 0.0)) + ee_contacts(E, ce.get("ee_st", 0.0))
        total += substeps * n_iters * per
    return total


def substep_bytes(V, E, B, n_iters=1):
    return predict(V) + n_iters * spring_pass(V, E, B) + update_vx(V)
