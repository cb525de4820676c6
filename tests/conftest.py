import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")


@pytest.fixture(scope="session")
def ref_counts():
    with open(os.path.join(GOLDEN, "ref_counts.json")) as f:
        return json.load(f)


def make_fixture_graph(spec):
    import subgraphmatchgpu_amd as sme

    if spec["kind"] == "rmat":
        g = sme.gen_rmat(spec["scale"], spec["edge_factor"], 0.57, 0.19, 0.19, spec["seed"])
    else:
        g = sme.gen_er(spec["n_vertices"], spec["n_edges"], spec["seed"])
    assert g.checksum() == spec["sha256"], "generator output changed; fixtures no longer apply"
    return g


@pytest.fixture(scope="session")
def config_counts():
    with open(os.path.join(GOLDEN, "config_counts.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def fixture_graphs(ref_counts):
    return {name: make_fixture_graph(spec) for name, spec in ref_counts["graphs"].items()}


def clique(k):
    import subgraphmatchgpu_amd as sme

    return sme.from_adjacency([[j for j in range(k) if j != i] for i in range(k)])


def member(g, offs, k):
    """Member k of a packed batch as a standalone graph (ids from 0)."""
    import numpy as np
    import subgraphmatchgpu_amd as sme

    a, b = int(offs[k]), int(offs[k + 1])
    rp = g.row_ptrs[a:b + 1].astype(np.uint64)
    cols = g.cols[int(rp[0]):int(rp[-1])].astype(np.int64) - a
    return sme.Graph(rp - rp[0], cols.astype(np.uint32))
