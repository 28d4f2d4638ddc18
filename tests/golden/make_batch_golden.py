"""Write FP64 oracle eigenvalues of the seeded shape collection.

    python tests/golden/make_batch_golden.py COUNT K

runs the oracle pipeline (numpy assembly as lapy/solver.py:106-194, splu +
eigsh shift-invert as lapy/solver.py:667-716) on every mesh of
lapy_amd.meshgen.shape_batch(count=COUNT) (icosphere levels 3-6, axes
U[0.5, 2]^3, radial noise U[0, 0.05], one seed per mesh) and stores the K
smallest eigenvalues of each in tests/golden/batch_{COUNT}_k{K}.json.
"""

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lapy_amd import meshgen  # noqa: E402
from oracle import fem  # noqa: E402


def main(count: int, k: int):
    meshes = meshgen.shape_batch(count=count)
    evals, sizes = [], []
    t0 = time.time()
    for v, t in meshes:
        a, b = fem.fem_tria(v, t)
        lam, _ = fem.eigs(a, b, k=k)
        evals.append(lam.tolist())
        sizes.append(int(a.shape[0]))
    out = {"count": count, "k": k, "n": sizes, "evals": evals, "oracle_seconds": time.time() - t0}
    path = os.path.join(HERE, f"batch_{count}_k{k}.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, out["oracle_seconds"])


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
