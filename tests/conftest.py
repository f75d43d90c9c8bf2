import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def O():
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "neighbor_golden.json")) as f:
        cases = json.load(f)
    for c in cases:
        c["x"] = np.array(c["position_bits"], np.uint32).view(np.float32).reshape(-1, 3)
        c["pairs_scalar"] = np.array(c["pairs_scalar"], np.int64).reshape(-1, 3)
        c["pairs_simd"] = np.array(c["pairs_simd"], np.int64).reshape(-1, 3)
        c["bruteforce"] = np.array(c["bruteforce"], np.int64).reshape(-1, 3)
    return cases


@pytest.fixture(scope="session")
def ctx(pkg):
    return pkg.default_context()


def pairs_to_rows(p):
    """(i, j, bits(d)) rows in emission order."""
    return np.stack([p["i"].astype(np.int64), p["j"].astype(np.int64),
                     np.ascontiguousarray(p["d"]).view(np.uint32).astype(np.int64)], axis=1)


def rows_canonical(a):
    a = np.asarray(a, np.int64).reshape(-1, 3)
    b = np.stack([np.minimum(a[:, 0], a[:, 1]), np.maximum(a[:, 0], a[:, 1]), a[:, 2]], axis=1)
    return b[np.lexsort((b[:, 2], b[:, 1], b[:, 0]))]


def bruteforce_np(x, r, pred):
    """All i < j pairs by the definition (float32 arithmetic, s folded as
    ((dx^2 + dy^2) + dz^2)): pred 0 keeps s < fl(r*r), pred 1 sqrt(s) < r.
    Returns (i, j, bits(d)) rows in (i, j) order.  O(N^2): small N only."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    r = np.float32(r)
    r2 = np.float32(r * r)
    rows = []
    for i in range(len(x) - 1):
        d3 = x[i] - x[i + 1:]
        s = (d3[:, 0] * d3[:, 0] + d3[:, 1] * d3[:, 1]) + d3[:, 2] * d3[:, 2]
        d = np.sqrt(s)
        keep = (s < r2) if pred == 0 else (d < r)
        j = np.nonzero(keep)[0]
        if len(j):
            rows.append(np.stack([np.full(len(j), i, np.int64), (j + i + 1).astype(np.int64),
                                  d[j].view(np.uint32).astype(np.int64)], axis=1))
    return np.concatenate(rows) if rows else np.zeros((0, 3), np.int64)
