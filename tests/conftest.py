import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")
    config.addinivalue_line("markers", "slow: large CPU case (seconds)")


def _ensure_built():
    """Loads what is built; builds only what is missing."""
    from lbvh_amd import build as B
    from oracle import build_oracle as OB
    if not os.path.exists(OB.ORACLE_SO):
        OB.build_oracle()
    if not os.path.exists(B.WORKLOAD_SO):
        B.build_workload()


_ensure_built()


def load_golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    data = np.load(path)
    cases = {}
    for key in data.files:
        case, field = key.split("/", 1)
        cases.setdefault(case, {})[field] = data[key]
    return cases


def golden_cases():
    out = []
    for group in ("small", "adversarial", "meshes"):
        for name, rec in load_golden(group).items():
            out.append((group + ":" + name, rec))
    return out


def f64_golden_cases():
    """builder<double> fixtures: (id, kind, prims, {leaf_order: node words u64[n-1, 8]})."""
    data = np.load(os.path.join(GOLDEN, "f64.npz"))
    cases = {}
    for key in data.files:
        kind, name, field = key.split("/", 2)
        cases.setdefault((kind, name), {})[field] = data[key]
    out = []
    for (kind, name), rec in sorted(cases.items()):
        out.append((kind + ":" + name, kind, rec["prims"],
                    {0: rec["l0_nodes"], 1: rec["l1_nodes"]}))
    return out


@pytest.fixture(scope="session")
def golden():
    return dict(golden_cases())


@pytest.fixture(scope="session")
def hashes():
    import json
    with open(os.path.join(GOLDEN, "hashes.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx():
    import lbvh_amd
    c = lbvh_amd.Context(0)
    yield c
    c.close()


def nodes_u32(nodes):
    return np.ascontiguousarray(nodes).view(np.uint32).reshape(-1, 8)
