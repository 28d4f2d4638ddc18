"""Mesh containers: the input contract LaPy's Solver reads from a mesh.

``TriaMesh`` / ``TetMesh`` keep ``.v`` (n, 3) and ``.t`` (m, 3|4) with the
reference's validation (lapy/tria_mesh.py:238-276, lapy/tet_mesh.py:51-59)
and the OFF / VTK-ASCII readers whose float32 vertex semantics the element
math depends on (lapy/_tria_io.py:50-227, lapy/_tet_io.py:131-224).  The
Solver dispatches on the class *name* (lapy/solver.py:72,95), so reference
mesh objects work as well.  ``avg_edge_length`` runs on the GPU.
"""

from __future__ import annotations

import numpy as np

from . import _lib, curvature


class TriaMesh:
    """Triangle mesh (vertices ``v`` (n, 3), triangles ``t`` (m, 3))."""

    def __init__(self, v, t, fsinfo=None):
        self.v = np.array(v)
        self.t = np.array(t)
        if self.v.ndim == 2 and self.v.shape[0] < self.v.shape[1]:
            self.v = self.v.T
        if self.t.ndim == 2 and self.t.shape[1] != 3 and self.t.shape[0] == 3:
            self.t = self.t.T
        if self.t.size == 0:
            raise ValueError("Mesh has no triangles (empty mesh)")
        if self.v.size == 0:
            raise ValueError("Mesh has no vertices (empty mesh)")
        vnum = np.max(self.v.shape)
        if np.max(self.t) >= vnum:
            raise ValueError("Max index exceeds number of vertices")
        if self.t.shape[1] != 3:
            raise ValueError("Triangles should have 3 vertices")
        if self.v.shape[1] == 2:
            self.v = np.column_stack([self.v, np.zeros(self.v.shape[0])])
            self._is_2d = True
        elif self.v.shape[1] == 3:
            self._is_2d = False
        else:
            raise ValueError("Vertices should have 2 or 3 coordinates")
        self.fsinfo = fsinfo

    def is_2d(self) -> bool:
        return self._is_2d

    @classmethod
    def read_off(cls, filename: str) -> "TriaMesh":
        v, t = _read_off(filename)
        return cls(v, t)

    @classmethod
    def read_vtk(cls, filename: str) -> "TriaMesh":
        v, t = _read_vtk(filename, 3)
        return cls(v, t)

    def avg_edge_length(self) -> float:
        """Mean length of the mesh edges (lapy/tria_mesh.py:735-748), on the GPU,
        in the vertex dtype as numpy computes it."""
        return _vertex_scalar(self, device_mesh(self).avg_edge_length())

    def area(self) -> float:
        v1, v2, v3 = self.v[self.t[:, 0]], self.v[self.t[:, 1]], self.v[self.t[:, 2]]
        return float(0.5 * np.linalg.norm(np.cross(v2 - v1, v3 - v1), axis=1).sum())

    def tria_normals(self) -> np.ndarray:
        """Unit triangle normals (lapy/tria_mesh.py:750-773)."""
        return curvature.tria_normals(self.v, self.t)

    def vertex_normals(self) -> np.ndarray:
        """Area-weighted unit vertex normals (lapy/tria_mesh.py:775-817)."""
        return curvature.vertex_normals(self.v, self.t)

    def curvature(self, smoothit: int = 3):
        """Per-vertex principal curvatures and directions
        (lapy/tria_mesh.py:1078-1219): ``u_min, u_max, c_min, c_max, c_mean,
        c_gauss, normals``, on the GPU."""
        return curvature.curvature(self.v, self.t, smoothit, dm=device_mesh(self))

    def curvature_tria(self, smoothit: int = 3):
        """Per-triangle principal directions and curvatures, the input of
        ``Solver(aniso=...)`` (lapy/tria_mesh.py:1221-1279), on the GPU."""
        return curvature.curvature_tria(self.v, self.t, smoothit, dm=device_mesh(self))


class TetMesh:
    """Tetrahedral mesh (vertices ``v`` (n, 3), tetrahedra ``t`` (m, 4))."""

    def __init__(self, v, t):
        self.v = np.array(v)
        self.t = np.array(t)
        vnum = self.v.shape[0]
        if np.max(self.t) >= vnum:
            raise ValueError("Max index exceeds number of vertices")

    @classmethod
    def read_vtk(cls, filename: str) -> "TetMesh":
        v, t = _read_vtk(filename, 4)
        return cls(v, t)

    def avg_edge_length(self) -> float:
        """Mean length of the mesh edges (lapy/tet_mesh.py:182-195), on the GPU,
        in the vertex dtype as numpy computes it."""
        return _vertex_scalar(self, device_mesh(self).avg_edge_length())


def _vertex_scalar(mesh, value):
    dt = np.asarray(mesh.v).dtype
    return np.float32(value) if dt == np.float32 else np.float64(value)


# ---------------------------------------------------------------------------
# device mesh cache: one lh_mesh per (vertices, elements) content
# ---------------------------------------------------------------------------

def _fingerprint(v, t):
    import hashlib

    h = hashlib.blake2b(digest_size=16)
    for a in (np.ascontiguousarray(v), np.ascontiguousarray(t)):
        h.update(str((a.dtype.str, a.shape)).encode())
        h.update(memoryview(a).cast("B"))
    return h.hexdigest()


def device_mesh(geometry, ctx=None) -> "_lib.DeviceMesh":
    """Device copy of ``geometry.v`` / ``geometry.t``, cached on the object
    while their contents are unchanged (the topology -- incidence lists and
    CSR pattern -- is built once per cached device mesh)."""
    ctx = ctx or _lib.default_context()
    v = np.asarray(geometry.v)
    t = np.asarray(geometry.t)
    if v.dtype not in (np.float32, np.float64):
        v = v.astype(np.float64)
    key = _fingerprint(v, t)
    cache = getattr(geometry, "_lapy_amd_dev", None)
    if cache is not None and cache[0] == key and cache[1].ctx is ctx:
        return cache[1]
    dm = _lib.DeviceMesh(ctx, v, t)
    try:
        geometry._lapy_amd_dev = (key, dm)
    except AttributeError:
        pass
    return dm


# ---------------------------------------------------------------------------
# readers (float32 vertices, as the reference's readers return)
# ---------------------------------------------------------------------------

def _read_off(filename):
    with open(filename) as f:
        line = f.readline()
        while line[0] == "#":
            line = f.readline()
        if not line.startswith("OFF"):
            raise OSError("[OFF keyword not found] --> FAILED\n")
        larr = f.readline().split()
        pnum, tnum = int(larr[0]), int(larr[1])
        v = np.fromfile(f, "float32", 3 * pnum, " ").reshape(pnum, 3)
        t = np.fromfile(f, "int", 4 * tnum, " ").reshape(tnum, 4)
    if np.amax(t[:, 0]) != 3:
        raise OSError("[no triangle data] --> FAILED\n")
    return v, t[:, 1:]


def _read_vtk(filename, nodes):
    with open(filename) as f:
        line = f.readline()
        while line[0] == "#":
            line = f.readline()
        count = 0
        while count < 5 and not line.startswith("ASCII"):
            line = f.readline()
            count += 1
        if not line.startswith("ASCII"):
            raise OSError("[ASCII keyword not found] --> FAILED\n")
        line = f.readline()
        if not line.startswith("DATASET POLYDATA") and not line.startswith("DATASET UNSTRUCTURED_GRID"):
            raise OSError(f"[read: {line} expected DATASET POLYDATA or DATASET UNSTRUCTURED_GRID] --> FAILED\n")
        larr = f.readline().split()
        if larr[0] != "POINTS" or (larr[2] != "float" and larr[2] != "double"):
            raise OSError("[expected POINTS # float or POINTS # double ] --> FAILED\n")
        pnum = int(larr[1])
        v = np.fromfile(f, "float32", 3 * pnum, " ").reshape(pnum, 3)
        larr = f.readline().split()
        if larr[0] not in ("POLYGONS", "CELLS"):
            raise OSError(f"[read: {larr} expected POLYGONS or CELLS] --> FAILED\n")
        tnum, ttnum = int(larr[1]), int(larr[2])
        if ttnum != (nodes + 1) * tnum:
            raise OSError(f"[having: {ttnum / tnum} data per element, expected {nodes}+1] --> FAILED\n")
        t = np.fromfile(f, "int", ttnum, " ").reshape(tnum, nodes + 1)
    if t[tnum - 1][0] != nodes:
        raise OSError("[unexpected element type] --> FAILED\n")
    return v, np.delete(t, 0, 1)
