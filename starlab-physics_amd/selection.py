"""Scene files of the reference's selection tool and animation setup, without
the GUI: pinned/handle vertex groups (animation/handle.json), handle animations
(framework/animation/animation.json) and sewing pairs (animation/sewing.csv).

Mirrors framework/utilities/selection_tool.py:105-226 (get_selection_array,
sewing_selection, pop_sewing, remove_all_sewing, export_selection,
import_selection), main.py:122-141 (load_animation) and XPBD.py:405-417
(load_sewing_pairs).  A selection array holds, per dynamic vertex, 0 or the
number (1..num_groups) of the group that selected it; Solver.set_fixed_vertices
pins every vertex whose entry is >= 1 (XPBD.py:389-395).
"""
from __future__ import annotations

import csv
import json
import os

import numpy as np

NUM_GROUPS = 4          # selection_tool.py num_maxCounter (groups 1..4)


def load_handles(path: str) -> dict:
    """import_selection (selection_tool.py:211-215): {group: [vertex ids]}."""
    with open(path, encoding="utf-8") as f:
        raw = json.load(f)
    return {int(k): [int(v) for v in vals] for k, vals in raw.items()}


def handles_from_selection(is_selected, num_groups: int = NUM_GROUPS) -> dict:
    """export_selection's grouping (selection_tool.py:176-195): the ids whose entry
    equals each group number, ascending."""
    sel = np.asarray(is_selected)
    return {g: [int(i) for i in np.argwhere(sel == g)[:, 0]] for g in range(1, num_groups + 1)}


def save_handles(path: str, groups) -> None:
    """export_selection (selection_tool.py:176-199): JSON with groups "1".."4",
    indented by 4 as the reference writes it.  `groups`: a dict or a selection
    array."""
    if not isinstance(groups, dict):
        groups = handles_from_selection(groups)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w", encoding="utf-8") as f:
        json.dump({int(k): list(v) for k, v in groups.items()}, f, ensure_ascii=False, indent=4)


def selection_array(groups: dict, num_verts: int) -> np.ndarray:
    """import_selection's is_selected (selection_tool.py:217-224): 0, or the group
    number of each vertex (later groups overwrite earlier ones)."""
    sel = np.zeros(num_verts, np.int32)
    for g in sorted(groups):
        idx = np.asarray(groups[g], np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= num_verts):
            raise IndexError(f"handle group {g} names a vertex outside [0, {num_verts})")
        sel[idx] = g
    return sel


def load_animation(path: str, num_groups: int = NUM_GROUPS) -> dict:
    """main.py:122-141 load_animation: per group a list of fragments
    [vx, vy, vz, rx, ry, rz, frame]; a group's list must hold a multiple of 7
    numbers."""
    with open(path, encoding="utf-8") as f:
        raw = {int(k): v for k, v in json.load(f).items()}
    out = {g: [] for g in range(1, num_groups + 1)}
    for g in range(1, num_groups + 1):
        vals = raw[g]
        if len(vals) % 7 != 0:
            raise ValueError(f"{g}th Animation SETTING ERROR!! ======")
        out[g] = [list(vals[7 * a:7 * a + 7]) for a in range(len(vals) // 7)]
    return out


def load_sewing_pairs(path: str) -> np.ndarray:
    """XPBD.py:405-417 load_sewing_pairs: int32 [n, 2] vertex pairs."""
    with open(path, "r", encoding="utf-8") as f:
        data = list(csv.reader(f))
    return np.array(data, dtype=np.int32).reshape(-1, 2)


def _write_pairs(path, pairs):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w", encoding="utf-8", newline="") as f:
        csv.writer(f).writerows([[int(a), int(b)] for a, b in pairs])


def add_sewing_pair(path: str, pair) -> bool:
    """sewing_selection (selection_tool.py:113-146): append the pair unless the
    same ordered pair is already there; returns whether it was added."""
    pair = [int(pair[0]), int(pair[1])]
    data = load_sewing_pairs(path).tolist() if os.path.exists(path) else []
    if pair in data:
        return False
    data.append(pair)
    _write_pairs(path, data)
    return True


def pop_sewing_pair(path: str):
    """pop_sewing (selection_tool.py:148-162): remove and return the last pair
    (None when there is none)."""
    data = load_sewing_pairs(path).tolist() if os.path.exists(path) else []
    if not data:
        return None
    last = data.pop()
    _write_pairs(path, data)
    return last


def remove_all_sewing(path: str) -> None:
    """remove_all_sewing (selection_tool.py:164-174)."""
    _write_pairs(path, [])
