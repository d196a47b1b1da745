"""starlab-physics on MI355X: the XPBD cloth step as gfx950 HIP kernels behind a
C ABI (include/slx.h), driven from Python with the reference's Solver surface.

Import with importlib (the directory name is not a Python identifier):

    slx = importlib.import_module("starlab-physics_amd")
    mesh = slx.Mesh(x=..., faces=...)
    sim = slx.Solver(mesh, None, dHat=1e-4, stiffness_stretch=1e6,
                     stiffness_bending=1e3, g=(0, -9.8, 0), dt=1/60)
    sim.forward(n_substeps=20)
"""
from . import scene, selection, topology  # noqa: F401
from ._abi import SlxError  # noqa: F401
from .mesh import Mesh  # noqa: F401
from .solver import Solver  # noqa: F401
