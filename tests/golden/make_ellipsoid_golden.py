"""Write FP64 oracle eigenvalues of the seeded ellipsoid workload.

    python tests/golden/make_ellipsoid_golden.py LEVEL K

runs the oracle pipeline (oracle/fem.py: numpy assembly as lapy/solver.py:
106-194, splu + eigsh shift-invert as lapy/solver.py:667-716) on
lapy_amd.meshgen.large_ellipsoid(level=LEVEL) and stores the K smallest
eigenvalues, eigenvector probes and timings in
tests/golden/ellipsoid_l{LEVEL}_k{K}.json.  Probes: the sign-normalised
eigenvectors at 256 seeded vertex indices, and per eigenvalue cluster the
basis-free projections P_g r = X_g X_g^T B r of 4 seeded normal vectors r
(seed 11) at the same vertices (oracle.metrics.projector_probes).
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


def main(level: int, k: int):
    v, t = meshgen.large_ellipsoid(level=level)
    t0 = time.time()
    a, b = fem.fem_tria(v, t)
    t1 = time.time()
    lam, vec = fem.eigs(a, b, k=k)
    t2 = time.time()
    idx = np.sort(np.random.default_rng(7).choice(len(v), size=min(256, len(v)), replace=False))
    vs = metrics.sign_normalise(vec)
    out = {
        "level": level, "k": k, "n": int(a.shape[0]), "nnz": int(a.nnz),
        "evals": lam.tolist(),
        "probe_index": idx.tolist(),
        "probe_vectors": vs[idx].tolist(),
        "cluster_probe_seed": 11,
        "cluster_probes": [{"cols": g, "values": p.tolist()}
                           for g, p in metrics.projector_probes(vec, b, lam, idx, seed=11)],
        "oracle_seconds": {"assembly": t1 - t0, "splu_eigsh": t2 - t1},
    }
    path = os.path.join(HERE, f"ellipsoid_l{level}_k{k}.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, out["oracle_seconds"])


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
