"""ctypes binding of libslx.so (the C ABI declared in include/slx.h).

The library is loaded from this package directory.  There is no fallback: if
libslx.so is missing, or a GPU entry point is called without a HIP device, the
call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SLX_LIB") or os.path.join(_HERE, "libslx.so")   # SLX_LIB: another build of the library
_LIB = None

I32, F32, PTR, STREAM = C.c_int32, C.c_float, C.c_void_p, C.c_void_p
I64 = C.c_int64

# name -> argument types (return type is always int32 status unless listed in _RET)
SIGNATURES = {
    "slx_abi_version": [],
    "slx_last_error": [],
    "slx_timer_enable": [I32],
    "slx_timer_read": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "slx_timer_names": [C.c_char_p, I32],
    "slx_timer_only": [C.c_char_p],
    "slx_compute_y": [I32, PTR, PTR, PTR, F32, F32, F32, F32, STREAM],
    "slx_spring_pass": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, F32, F32, STREAM],
    "slx_update_vx": [I32, PTR, PTR, PTR, PTR, F32, F32, STREAM],
    "slx_substeps_plain": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, I32, I32,
                           F32, F32, F32, F32, F32, F32, F32, STREAM],
    "slx_set_fixed": [I32, PTR, PTR, PTR, PTR, PTR, STREAM],
    "slx_set_spring_launch": [I32, I32],
    "slx_scan_ws_bytes": [I32],
    "slx_exclusive_scan": [I32, PTR, PTR, PTR, PTR, PTR, C.c_size_t, STREAM],
    "slx_radix_sort_ws_bytes": [I32],
    "slx_radix_sort_pairs": [I32, PTR, PTR, PTR, I32, PTR, C.c_size_t, STREAM],
    "slx_radix_sort_pairs_noalloc": [I32, PTR, PTR, PTR, I32, STREAM],
    "slx_face_aabbs": [I32, PTR, PTR, F32, PTR, PTR, STREAM],
    "slx_mesh_aabb_ws_bytes": [],
    "slx_mesh_aabb": [I32, PTR, F32, PTR, PTR, C.c_size_t, STREAM],
    "slx_cell_topology_ws_bytes": [],
    "slx_cell_topology": [PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, C.c_size_t, STREAM],
    "slx_refit_sched_ints": [I32],
    "slx_refit_schedule": [I32, PTR, PTR, PTR, PTR, PTR, STREAM],
    "slx_set_refit_shape": [I32, I32],
    "slx_cell_build_ws_bytes": [I32],
    "slx_cell_build": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR,
                       PTR, PTR, C.c_size_t, STREAM],
    "slx_tri_build_ws_bytes": [I32],
    "slx_tri_build": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR,
                      C.c_size_t, STREAM],
    "slx_tri_refit": [I32, PTR, PTR, F32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, STREAM],
    "slx_traverse": [I32, I32, PTR, F32, I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, I32, PTR, PTR, PTR, STREAM],
    "slx_traverse_edges": [I32, I32, PTR, PTR, F32, I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, I32, PTR, PTR,
                           PTR, PTR, PTR, STREAM],
    "slx_traverse_cached": [I32, I32, PTR, F32, I32, PTR, PTR, PTR, PTR, PTR, I32, PTR, PTR, PTR, PTR, STREAM],
    "slx_set_walk_stack": [I32],
    "slx_set_active_grid": [I32],
    "slx_tree_snapshot": [I32, PTR, PTR, PTR, STREAM],
    "slx_tree_displacement": [I32, PTR, PTR, PTR, STREAM],
    "slx_collision_ws_bytes": [I32, I32, I32],
    "slx_collision_pass": [I32, I32, PTR, PTR, PTR, PTR, PTR, PTR, F32, F32, I32, PTR, PTR, PTR, PTR, PTR,
                           PTR, PTR, I32, PTR, PTR, I32, PTR, PTR, C.c_size_t, STREAM],
    "slx_collision_pass_ee": [I32, I32, PTR, PTR, PTR, PTR, PTR, PTR, F32, F32, I32, PTR, PTR, PTR, PTR, PTR,
                              PTR, PTR, I32, PTR, PTR, PTR, I32, PTR, PTR, C.c_size_t, STREAM],
    "slx_sewing_pass": [I32, PTR, PTR, PTR, PTR, PTR, PTR, F32, PTR, STREAM],
    "slx_compute_velocity": [I32, PTR, PTR, PTR, PTR, F32, F32, STREAM],
    "slx_update_x": [I32, PTR, PTR, F32, STREAM],
    "slx_velocity_pass": [I32, I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, F32, I32, PTR, PTR, PTR, F32, I32, PTR, PTR,
                          C.c_size_t, STREAM],
    "slx_velocity_pass_ee": [I32, I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, F32, I32, PTR, PTR, PTR, PTR, F32, I32,
                             PTR, PTR, C.c_size_t, STREAM],
    "slx_frame_ws_bytes": [PTR],
    "slx_frame": [PTR, STREAM],
    "slx_frame_ctx_create": [PTR],
    "slx_frame_ctx_destroy": [PTR],
    "slx_seg_build_ws_bytes": [I32, I32],
    "slx_seg_build": [I32, I32, I32, PTR, PTR, F32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR,
                      PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, C.c_size_t, STREAM],
    "slx_traverse_seg": [I32, PTR, PTR, PTR, F32, I32, PTR, PTR, PTR, I32, PTR, PTR, PTR, STREAM],
    "slx_mlp_num_params": [],
    "slx_features": [I32, PTR, PTR, PTR, PTR, PTR, PTR, F32, F32, PTR, STREAM],
    "slx_warm_start": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, F32, F32, PTR, PTR, PTR, PTR, F32, STREAM],
    "slx_mlp_grad": [I32, PTR, PTR, PTR, PTR, PTR, PTR, I32, PTR, PTR, PTR, STREAM],
    "slx_mlp_adam": [I32, PTR, PTR, PTR, PTR, PTR, F32, F32, F32, F32, I32, STREAM],
    "slx_mlp_to_bf16": [PTR, PTR, STREAM],
    "slx_max_update": [I32, PTR, PTR, PTR, STREAM],
    "slx_solve_reset": [I32, PTR, PTR, STREAM],
    "slx_implicit_passes": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, F32, I32, I32, F32, PTR, PTR, STREAM],
    "slx_greedy_edge_coloring": [I32, I32, PTR, I32, PTR, PTR, PTR],
    "slx_gs_pass": [I32, PTR, PTR, PTR, PTR, PTR, PTR, PTR, PTR, I32, F32, F32, STREAM],
}
_RET = {"slx_last_error": C.c_char_p, "slx_scan_ws_bytes": C.c_size_t,
        "slx_radix_sort_ws_bytes": C.c_size_t, "slx_mesh_aabb_ws_bytes": C.c_size_t,
        "slx_cell_topology_ws_bytes": C.c_size_t, "slx_cell_build_ws_bytes": C.c_size_t,
        "slx_tri_build_ws_bytes": C.c_size_t, "slx_collision_ws_bytes": C.c_size_t,
        "slx_frame_ws_bytes": C.c_size_t, "slx_seg_build_ws_bytes": C.c_size_t,
        "slx_refit_sched_ints": C.c_size_t}


class Tree(C.Structure):
    """slx_tree (include/slx.h)."""
    _fields_ = [("mode", I32), ("nf", I32), ("fmin4", PTR), ("fmax4", PTR), ("box6", PTR),
                ("codes", PTR), ("sorted_faces", PTR), ("child_a", PTR), ("child_b", PTR),
                ("parent", PTR), ("num", PTR), ("face_cell", PTR), ("prefix", PTR), ("leaf_l", PTR),
                ("leaf_r", PTR), ("nlo4", PTR), ("nhi4", PTR), ("n_inst", I32), ("vert_base", PTR),
                ("face_base", PTR), ("face_inst", PTR), ("vert_inst", PTR), ("sorted_inst", PTR),
                ("range_l", PTR), ("range_r", PTR), ("roots", PTR), ("sched", PTR),
                ("vert_order", PTR), ("quad4", PTR), ("x_snap", PTR), ("disp", PTR)]


class CandCache(C.Structure):
    """slx_cand_cache (include/slx.h)."""
    _fields_ = [("cap", I32), ("margin", F32), ("leaves", PTR), ("count", PTR), ("q0", PTR), ("redo", PTR)]


class EdgeSets(C.Structure):
    """slx_edge_sets (include/slx.h)."""
    _fields_ = [("ne", I32), ("edges_dy", PTR), ("edges_st", PTR), ("fe_dy", PTR), ("fe_st", PTR), ("cap", I32),
                ("c_st", PTR), ("n_st", PTR), ("c_dy", PTR), ("n_dy", PTR), ("pair_cap", I32), ("pair_num", PTR),
                ("pair_cnt", PTR), ("pair_rec", PTR), ("order", PTR), ("redo", PTR), ("gate", PTR), ("live", PTR),
                ("nes", I32)]


class FrameDesc(C.Structure):
    """slx_frame_desc (include/slx.h)."""
    _fields_ = [("nv", I32), ("x4", PTR), ("v4", PTR), ("y4", PTR), ("y4_alt", PTR), ("fixed", PTR),
                ("m_inv", PTR), ("faces", PTR), ("adj_slice_off", PTR), ("adj_slice_width", PTR),
                ("adj_entries", PTR), ("nvs", I32), ("xs4", PTR), ("faces_st", PTR),
                ("dt", F32), ("gx", F32), ("gy", F32), ("gz", F32), ("k_s", F32), ("k_b", F32),
                ("damping", F32), ("padding", F32), ("dhat", F32), ("k_col", F32),
                ("n_substeps", I32), ("n_iters", I32), ("collision", I32), ("refit_every_substep", I32),
                ("rebuild_static", I32), ("cap", I32), ("rec_cap", I32), ("method", I32), ("gs_edges", PTR),
                ("gs_color_off", PTR), ("n_colors", I32), ("dyn", Tree), ("st", Tree),
                ("cand", PTR), ("cand_n", PTR), ("pair_num", PTR), ("status", PTR), ("ws", PTR),
                ("ws_bytes", C.c_size_t), ("velocity", I32), ("mu", F32), ("vs4", PTR), ("m_inv_st", PTR),
                ("pair_cap", I32), ("pair_cnt", PTR), ("pair_rec", PTR), ("cache", CandCache * 3), ("cache_every", I32),
                ("n_sewn", I32), ("sew_verts", PTR), ("sew_off", PTR), ("sew_list", PTR), ("sew_pairs", PTR),
                ("k_sew", F32), ("sew_ws", PTR), ("ee", EdgeSets), ("ctx", PTR), ("spring_block", I32),
                ("spring_vpt", I32), ("walk_stack", I32), ("active_grid", I32), ("refit_sub", I32),
                ("refit_top", I32), ("walk_wave", I32), ("own_faces", PTR),
                ("cache_every_static", I32)]


class SlxError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SlxError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                           "python starlab-physics_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, C.c_int)
        _LIB = L
    return _LIB


def exported_symbols():
    return list(SIGNATURES)


def last_error() -> str:
    msg = lib().slx_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise SlxError(f"{name} failed ({rc}): {last_error()}")


STATUS_BITS = {1: "candidate list overflow (contacts past the capacity dropped)",
               2: "traversal stack overflow",
               4: "face centre outside the 32^3 cell grid",
               8: "contact records per vertex exceeded (deltas dropped)",
               16: "contacting pairs per query exceeded (dropped)"}


def status_flags(st: int) -> str:
    """The SLX_STATUS_* bits of a status word, named."""
    names = [v for b, v in STATUS_BITS.items() if st & b]
    if st & ~sum(STATUS_BITS):
        names.append(f"unknown bits {st & ~sum(STATUS_BITS):#x}")
    return "; ".join(names)


def ptr(t) -> int:
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def timer_enable(max_launches: int) -> None:
    call("slx_timer_enable", max_launches)


def timer_only(name=None) -> None:
    call("slx_timer_only", name.encode() if name else None)


def timer_read(name: str):
    """(total_ms, count) of the timed launches of kernel `name`."""
    tot, cnt = C.c_double(0.0), C.c_int(0)
    call("slx_timer_read", name.encode(), C.byref(tot), C.byref(cnt))
    return tot.value, cnt.value


def timer_names():
    buf = C.create_string_buffer(4096)
    call("slx_timer_names", buf, len(buf))
    v = buf.value.decode()
    return v.split(",") if v else []


def greedy_coloring(nv: int, edges: np.ndarray, max_colors: int = 64):
    """Host greedy edge colouring in libslx (slx_greedy_edge_coloring)."""
    e = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
    ne = len(e)
    color = np.zeros(max(ne, 1), np.int32)
    order = np.zeros(max(ne, 1), np.int32)
    offsets = np.zeros(max_colors + 1, np.int32)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    nc = lib().slx_greedy_edge_coloring(nv, ne, P(e), max_colors, P(color), P(order), P(offsets))
    if nc < 0:
        raise SlxError(f"slx_greedy_edge_coloring failed: {last_error()}")
    return color[:ne], order[:ne], offsets[:nc + 1].copy()
