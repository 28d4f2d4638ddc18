"""Write FP64 oracle results of the seeded tetrahedral workload.

    python tests/golden/make_volume_golden.py CELLS K

runs the oracle pipeline on lapy_amd.meshgen.volume_cube(cells=CELLS) (unit
cube, 6 tetrahedra per cell in the split of the reference's cubeTetra.vtk,
interior vertices jittered U(-0.15 h, 0.15 h)):
  * eigenpairs: numpy tetrahedral assembly (lapy/solver.py:380-533), splu +
    eigsh shift-invert (lapy/solver.py:667-716), the K smallest eigenvalues
    and sign-normalised eigenvector probes at 256 seeded vertices, plus per
    eigenvalue cluster the basis-free projections P_g r = X_g X_g^T B r of 4
    seeded normal vectors r (seed 11) at the same vertices;
  * heat diffusion from vertex 0 with t = 1e-3 (lapy/heat.py:114-232, i.e.
    m = 1e-3 / avg_edge_length^2), values at the same probe vertices.
Output: tests/golden/volume_c{CELLS}_k{K}.json.
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lapy_amd import meshgen  # noqa: E402
from oracle import fem, metrics  # noqa: E402


def main(cells: int, k: int):
    v, t = meshgen.volume_cube(cells=cells)
    t0 = time.time()
    a, b = fem.fem_tetra(v, t)
    t1 = time.time()
    lam, vec = fem.eigs(a, b, k=k)
    t2 = time.time()
    idx = np.sort(np.random.default_rng(7).choice(len(v), size=min(256, len(v)), replace=False))
    vs = metrics.sign_normalise(vec)
    ael = float(fem.avg_edge_length(v, t))
    m = 1e-3 / (ael * ael)
    t3 = time.time()
    u = fem.diffusion(v, t, 0, m=m)
    t4 = time.time()
    out = {
        "cells": cells, "k": k, "n": int(a.shape[0]), "nnz": int(a.nnz),
        "evals": lam.tolist(),
        "probe_index": idx.tolist(),
        "probe_vectors": vs[idx].tolist(),
        "cluster_probe_seed": 11,
        "cluster_probes": [{"cols": g, "values": p.tolist()}
                           for g, p in metrics.projector_probes(vec, b, lam, idx, seed=11)],
        "avg_edge_length": ael,
        "heat_m": m,
        "heat_probe_values": u[idx].tolist(),
        "heat_sum": float(u.sum()),
        "oracle_seconds": {"assembly": t1 - t0, "splu_eigsh": t2 - t1, "heat_splu": t4 - t3},
    }
    path = os.path.join(HERE, f"volume_c{cells}_k{k}.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, out["oracle_seconds"])


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
