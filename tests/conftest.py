import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DATA = os.path.join(ROOT, "tests", "golden", "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblapyhip.so")
    config.addinivalue_line("markers", "slow: longer-running test")


def data_path(name):
    return os.path.join(DATA, name)


def read_ev(path):
    """Minimal reader of the ShapeDNA .ev text format: (eigenvalues, vectors (n, k))."""
    import re

    with open(path) as f:
        txt = f.read()
    ev_block = txt.split("Eigenvalues:")[1].split("}")[0].split("{")[1]
    evals = np.array([float(x) for x in ev_block.split(";")])
    vecs = None
    if "Eigenvectors:" in txt:
        rest = txt.split("Eigenvectors:")[1]
        n, k = (int(x) for x in re.search(r"sizes:\s*(\d+)\s+(\d+)", rest).groups())
        groups = re.findall(r"\(([^)]*)\)", rest)
        vecs = np.array([[float(x) for x in g.split(",")] for g in groups[:k]]).T
        assert vecs.shape == (n, k)
    return evals, vecs


@pytest.fixture(scope="session")
def expected():
    import json

    with open(data_path("expected_outcomes.json")) as f:
        return json.load(f)["expected_outcomes"]
