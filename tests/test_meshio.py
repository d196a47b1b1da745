"""Mesh and scene file formats (CPU): OBJ reading as libigl does for the
reference's own assets, the concatenation of framework/meshio/concat.py with
its vertex offsets, per-frame export paths (meshTaichiWrapper.py:369-386), the
selection tool's handle JSON (byte-identical round trip of the reference's own
files), handle animations and sewing pairs.  Fixtures: data files copied from
the reference by tests/golden/make_meshio_fixtures.py."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshio")


def test_read_reference_objs(slx):
    m = slx.scene.read_obj(os.path.join(GOLD, "plane_8.obj"))
    assert m.x.shape == (81, 3) and m.faces.shape == (128, 3)
    assert m.x.dtype == np.float32 and m.faces.min() == 0 and m.faces.max() == 80
    np.testing.assert_array_equal(m.x[0], [-1.0, 0.0, 1.0])
    np.testing.assert_array_equal(m.faces[0], [9, 58, 57])         # "f 10 59 58", 1-based
    sq = slx.scene.read_obj(os.path.join(GOLD, "square.obj"), dtype=np.float64)
    assert sq.x.dtype == np.float64
    np.testing.assert_array_equal(sq.faces, [[0, 1, 3], [3, 2, 0]])


def test_read_obj_face_forms(slx, tmp_path):
    """v/vt/vn and v/vt faces (BasicTShirt.obj, dress.obj), trailing blanks,
    polygons (fan), negative indices."""
    p = tmp_path / "forms.obj"
    p.write_text("o x\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\ns 0\n"
                 "f 1/1/1 2/1/1 3/1/1\nf 1/1 3/1 4/1 \nf -4 -3 -2 -1\n")
    m = slx.scene.read_obj(str(p))
    np.testing.assert_array_equal(m.faces, [[0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 2, 3]])


def test_concat_mesh_offsets_and_transform(slx, tmp_path):
    """concat.py:11-58: per model scale then translate (doubles), faces offset by
    the running vertex count, offsets [0, n0, n0+n1, ..., total]; the written OBJ
    reads back to the same doubles."""
    names = ["plane_8.obj", "square.obj", "plane_8.obj"]
    trans = [[0.0, 7.0, 0.0], [0.5, 7.5, -0.25], [0.0, 8.0, 0.0]]
    scales = [3.4, 3.0, 0.1]
    offs = slx.scene.concat_mesh("concat.obj", GOLD, names, trans, scales, out_dir=str(tmp_path))
    assert offs == [0, 81, 85, 166]
    got = slx.scene.read_obj(str(tmp_path / "concat.obj"), dtype=np.float64)
    for k, name in enumerate(names):
        m = slx.scene.read_obj(os.path.join(GOLD, name), dtype=np.float64)
        a, b = offs[k], offs[k + 1]
        np.testing.assert_array_equal(got.x[a:b], scales[k] * m.x + np.asarray(trans[k]))
        fa = sum(len(slx.scene.read_obj(os.path.join(GOLD, n)).faces) for n in names[:k])
        np.testing.assert_array_equal(got.faces[fa:fa + len(m.faces)], m.faces + a)
    # the reference scene then loads the concatenation as one dynamic mesh (Scenes/concat_test.py:41)
    mesh = slx.Mesh(str(tmp_path), "concat.obj", offsets=offs, device=None)
    assert mesh.num_verts == 166 and mesh.offsets == offs
    assert mesh.num_faces == 128 + 2 + 128


def test_export_paths(slx, tmp_path):
    """meshTaichiWrapper.py:369-386: <root>/<scene>/Mesh_ID_/Mesh_obj_<frame>.obj
    (StaticMesh_ID_ for static meshes), x as it is."""
    import torch
    if not torch.cuda.is_available():
        g = slx.scene.grid_cloth(3)
        m = slx.Mesh(x=g.x, faces=g.faces, device=None)
        m.positions = lambda: m.x0_np          # host-only mesh: export its rest positions
    else:
        g = slx.scene.grid_cloth(3)
        m = slx.Mesh(x=g.x, faces=g.faces)
    p = m.export("scene_a", 7, root=str(tmp_path))
    assert p == os.path.join(str(tmp_path), "scene_a", "Mesh_ID_", "Mesh_obj_7.obj")
    back = slx.scene.read_obj(p)
    np.testing.assert_array_equal(back.x, m.positions())
    np.testing.assert_array_equal(back.faces, m.fid_np)
    p2 = m.export("scene_a", 0, is_static=True, root=str(tmp_path))
    assert os.path.dirname(p2).endswith("StaticMesh_ID_")


@pytest.mark.parametrize("name", ["handle.json", "handle_thin_shell_twist.json", "handle_cylinder_crossing.json"])
def test_handle_json_round_trip_is_byte_identical(slx, tmp_path, name):
    src = os.path.join(GOLD, name)
    groups = slx.selection.load_handles(src)
    assert sorted(groups) == [1, 2, 3, 4]
    nv = max(max(v) for v in groups.values() if v) + 10
    sel = slx.selection.selection_array(groups, nv)
    for g, ids in groups.items():
        assert (sel[ids] == g).all()
    assert (sel > 0).sum() == len(set(i for v in groups.values() for i in v))
    out = tmp_path / "handle.json"
    slx.selection.save_handles(str(out), sel)                     # from the selection array, as export_selection
    assert out.read_bytes() == open(src, "rb").read()
    slx.selection.save_handles(str(out), groups)
    assert out.read_bytes() == open(src, "rb").read()


def test_handles_pin_through_set_fixed_semantics(slx):
    """selection >= 1 -> pinned (fixed 0), else free (XPBD.py:389-395)."""
    groups = slx.selection.load_handles(os.path.join(GOLD, "handle_thin_shell_twist.json"))
    sel = slx.selection.selection_array(groups, 5000)
    fixed = np.where(sel >= 1, 0.0, 1.0)
    assert fixed.sum() == 5000 - 132
    with pytest.raises(IndexError):
        slx.selection.selection_array({1: [5]}, 5)


def test_animation_json(slx, tmp_path):
    a = slx.selection.load_animation(os.path.join(GOLD, "animation.json"))
    assert a == {1: [[0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 200]], 2: [[0.0, 1.0, 0.0, -0.0, 0.0, 0.0, 200]], 3: [], 4: []}
    b = slx.selection.load_animation(os.path.join(GOLD, "animation_thin_shell_twist.json"))
    assert b[1] == [[0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 5000]] and b[2][0][3] == -0.5
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"1": [0.0] * 8, "2": [], "3": [], "4": []}))
    with pytest.raises(ValueError, match="1th Animation SETTING ERROR"):
        slx.selection.load_animation(str(bad))


def test_sewing_pairs_file(slx, tmp_path):
    p = str(tmp_path / "animation" / "sewing.csv")
    assert slx.selection.add_sewing_pair(p, (3, 7))
    assert slx.selection.add_sewing_pair(p, (8, 2))
    assert not slx.selection.add_sewing_pair(p, (3, 7))            # duplicate
    assert slx.selection.add_sewing_pair(p, (7, 3))                # the other order is another pair
    np.testing.assert_array_equal(slx.selection.load_sewing_pairs(p), [[3, 7], [8, 2], [7, 3]])
    assert open(p, newline="").read() == "3,7\r\n8,2\r\n7,3\r\n"   # csv.writer rows, as the reference writes
    assert slx.selection.pop_sewing_pair(p) == [7, 3]
    slx.selection.remove_all_sewing(p)
    assert slx.selection.load_sewing_pairs(p).shape == (0, 2)
    assert slx.selection.pop_sewing_pair(p) is None
