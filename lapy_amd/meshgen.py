"""Seeded synthetic meshes of the benchmark workloads.

* ``icosphere(level)``: icosahedron refined ``level`` times by 1-to-4 midpoint
  subdivision (the route of lapy/tria_mesh.py:1452-1487) with the new vertices
  projected to the unit sphere: V = 10*4^L + 2, F = 20*4^L.
* ``perturbed_sphere``: ellipsoidal scaling and radial noise.
* ``tet_cube(cells)``: unit cube, ``cells``^3 cubes split into 6 tetrahedra
  each with the split and ordering of the reference fixture
  ``data/cubeTetra.vtk`` (vertex index x fastest, cells z fastest), optional
  interior jitter.
"""

from __future__ import annotations

import numpy as np

SEED_BASE = 20260101


def icosahedron():
    p = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([
        [-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
        [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
        [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1],
    ], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    t = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    return v, t


def subdivide(v, t, project=True):
    """One 1-to-4 midpoint refinement; new vertex per unique edge (sorted)."""
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]], axis=0)
    e.sort(axis=1)
    key = e[:, 0] * (len(v) + 1) + e[:, 1]
    uniq, inv = np.unique(key, return_inverse=True)
    ei = np.stack([uniq // (len(v) + 1), uniq % (len(v) + 1)], axis=1)
    mid = 0.5 * (v[ei[:, 0]] + v[ei[:, 1]])
    if project:
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    nv = len(v)
    m = len(t)
    e01 = nv + inv[:m]
    e12 = nv + inv[m:2 * m]
    e20 = nv + inv[2 * m:]
    t0, t1, t2 = t[:, 0], t[:, 1], t[:, 2]
    tn = np.stack([
        np.stack([t0, e01, e20], 1),
        np.stack([t1, e12, e01], 1),
        np.stack([t2, e20, e12], 1),
        np.stack([e01, e12, e20], 1),
    ], axis=1).reshape(-1, 3)
    return np.concatenate([v, mid], axis=0), tn


def icosphere(level: int):
    v, t = icosahedron()
    for _ in range(level):
        v, t = subdivide(v, t, project=True)
    return v, t


def perturbed_sphere(level: int, rng, axes=None, sigma=0.0):
    """Icosphere with radial noise ``(1 + sigma N(0,1))`` and axis scaling."""
    v, t = icosphere(level)
    if sigma:
        v = v * (1.0 + sigma * rng.standard_normal(len(v)))[:, None]
    if axes is not None:
        v = v * np.asarray(axes, dtype=np.float64)[None, :]
    return v, t


def small_sphere(seed: int = 0, level: int = 5):
    """Icosphere level 5 (10,242 vertices), radial noise 0.02."""
    rng = np.random.default_rng(SEED_BASE + seed)
    return perturbed_sphere(level, rng, sigma=0.02)


def large_ellipsoid(seed: int = 0, level: int = 8):
    """Icosphere level 8 (655,362 vertices), axes U[0.5, 2]^3, noise 0.01."""
    rng = np.random.default_rng(SEED_BASE + seed)
    axes = rng.uniform(0.5, 2.0, size=3)
    return perturbed_sphere(level, rng, axes=axes, sigma=0.01)


def shape_mesh(i: int, seed: int = 0, levels=(3, 4, 5, 6)):
    """Mesh ``i`` of a seeded collection: level ~ U{levels}, axes U[0.5,2]^3,
    noise sigma ~ U[0, 0.05]."""
    rng = np.random.default_rng(SEED_BASE + seed + 1000 + i)
    lev = int(rng.choice(levels))
    axes = rng.uniform(0.5, 2.0, size=3)
    sig = rng.uniform(0.0, 0.05)
    return perturbed_sphere(lev, rng, axes=axes, sigma=sig)


def shape_batch(count: int = 256, seed: int = 0, levels=(3, 4, 5, 6)):
    """``count`` ellipsoids of uneven size (``shape_mesh`` 0 .. count-1)."""
    return [shape_mesh(i, seed, levels) for i in range(count)]


# the learned preconditioner's collections: 512 training meshes of levels
# 4-7 and 64 held-out level-7 meshes (seeds disjoint from each other and from
# the 256-mesh batch)
TRAIN_SEED, HELD_OUT_SEED = 70_000, 90_000


def training_mesh(i: int):
    return shape_mesh(i, TRAIN_SEED, levels=(4, 5, 6, 7))


def held_out_mesh(i: int):
    return shape_mesh(i, HELD_OUT_SEED, levels=(7,))


def tet_cube(cells: int, jitter: float = 0.0, seed: int = 0):
    """Unit cube with ``cells``^3 cubes of 6 tetrahedra (cubeTetra.vtk split)."""
    n1 = cells + 1
    h = 1.0 / cells
    g = np.arange(n1, dtype=np.float64) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)  # index = x + n1*y + n1^2*z
    if jitter:
        rng = np.random.default_rng(SEED_BASE + seed)
        interior = np.all((v > 0.5 * h) & (v < 1.0 - 0.5 * h), axis=1)
        d = rng.uniform(-jitter * h, jitter * h, size=v.shape)
        v = v + d * interior[:, None]
    ci = np.arange(cells)
    cx, cy, cz = np.meshgrid(ci, ci, ci, indexing="ij")  # z fastest, x slowest
    base = (cx + n1 * cy + n1 * n1 * cz).ravel()
    c000 = base
    c100 = base + 1
    c010 = base + n1
    c110 = base + 1 + n1
    c001 = base + n1 * n1
    c101 = c001 + 1
    c011 = c001 + n1
    c111 = c001 + 1 + n1
    tets = np.stack([
        np.stack([c000, c100, c110, c101], 1),
        np.stack([c000, c110, c001, c101], 1),
        np.stack([c110, c001, c101, c111], 1),
        np.stack([c001, c010, c111, c011], 1),
        np.stack([c000, c110, c001, c010], 1),
        np.stack([c110, c001, c010, c111], 1),
    ], axis=1).reshape(-1, 4)
    return v, tets.astype(np.int64)


def volume_cube(seed: int = 0, cells: int = 64):
    """64^3-cell tet cube (274,625 vertices), interior jitter U(-0.15h, 0.15h)."""
    return tet_cube(cells, jitter=0.15, seed=seed)
