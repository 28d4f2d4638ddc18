"""lapy_amd: LaPy's FEM hot path (assembly, eigensolve, heat / Poisson) on
AMD Instinct MI355X (gfx950), behind LaPy's own Python API.

    from lapy_amd import TriaMesh, TetMesh, Solver, heat, shapedna, io
"""

from . import diffgeo, heat, io, shapedna  # noqa: F401
from .mesh import TetMesh, TriaMesh  # noqa: F401
from .solver import Solver  # noqa: F401

__version__ = "0.1.0"
