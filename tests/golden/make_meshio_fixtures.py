"""Copies the small data files of the reference that pin the mesh and scene file
formats into tests/golden/meshio/ (data only: OBJ meshes shipped as assets, the
handle/pin selections and handle animations written by the reference's own
selection tool).  Run in a checkout that has the reference at /root/reference:

    python tests/golden/make_meshio_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "models/OBJ/plane_8.obj": "plane_8.obj",
    "models/OBJ/square.obj": "square.obj",
    "models/OBJ/curved_plane.obj": "curved_plane.obj",
    "framework/animation/handle.json": "handle.json",
    "framework/animation/thin_shell_twist/handle.json": "handle_thin_shell_twist.json",
    "framework/animation/cylinder_crossing/handle.json": "handle_cylinder_crossing.json",
    "framework/animation/animation.json": "animation.json",
    "framework/animation/thin_shell_twist/animation.json": "animation_thin_shell_twist.json",
}

if __name__ == "__main__":
    out = os.path.join(HERE, "meshio")
    os.makedirs(out, exist_ok=True)
    for src, dst in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(out, dst))
        print(dst)
