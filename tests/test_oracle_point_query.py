"""Oracle checks of the point-query traversal (TreeData + neighbor_traverse,
BVHTraverse.jl:500-543, 1057-1092, 1125-1234; the root-started form of
gpuneighbor_traverse!, ext/NaiveKA.jl:352-452), TreeData's store and
quantized_positions! (BVHTraverse.jl:290-314).

Pins:
  * tests/golden/point_query_golden.json, written from the definitions with no
    tree walk (leaf-box brute force, Python's stable sort for ranks); the
    cases include position8 of test/BVHTraverse.jl:165-169, points beyond the
    unit cube, tied and signed-zero coordinates, and two margin pairs inside
    the radius but on the boundary of the +-r box;
  * the reference test's facts for position8 at r = 10 (every atom paired,
    test/BVHTraverse.jl:181-198) and its set equality with the all-pairs list
    over radius, leaf-size and size sweeps (:203-280);
  * rank invariants: a permutation of 1..N per axis, monotone in the
    coordinate, ties ordered by the previous axis' order.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def pq_golden():
    with open(os.path.join(HERE, "golden", "point_query_golden.json")) as f:
        return json.load(f)


def _x(c):
    return np.array(c["position_bits"], np.uint32).view(np.float32)


def _rows(p):
    if len(p) == 0:
        return np.zeros((0, 3), np.int64)
    return np.stack([p["i"], p["j"], p["d"].view(np.uint32)], 1).astype(np.int64)


def _golden_rows(c, key):
    return np.array(c[key], np.int64).reshape(-1, 3)


def _unordered(p):
    return {(min(a, b), max(a, b)) for a, b in zip(p["i"].tolist(), p["j"].tolist())}


def _pair_keys(p):
    """Sorted int64 keys min(i,j) * 2^32 + max(i,j) of an unordered pair list."""
    i = np.minimum(p["i"], p["j"]).astype(np.int64)
    j = np.maximum(p["i"], p["j"]).astype(np.int64)
    return np.sort((i << 32) | j)


def _same_pairs(a, b):
    ka, kb = _pair_keys(a), _pair_keys(b)
    return ka.shape == kb.shape and (ka == kb).all()


@pytest.mark.parametrize("mode,key", [(0, "pairs_own_leaf"), (1, "pairs_root")])
def test_pairs_match_golden_in_order(O, pq_golden, mode, key):
    for c in pq_golden:
        p, _ = O.build_traverse_bvh(_x(c), c["apl"], c["r"], mode=mode)
        g = _golden_rows(c, key)
        assert _rows(p).shape == g.shape, c["name"]
        assert (_rows(p) == g).all(), c["name"]


def test_tree_and_store_match_golden(O, pq_golden):
    for c in pq_golden:
        x = _x(c)
        _, t = O.build_traverse_bvh(x, c["apl"], c["r"], mode=O.QUERY_ROOT)
        assert (t["perm"] == np.array(c["perm"])).all()
        for k, n in enumerate(c["nodes_root_kept"]):
            assert (t["nodes"][k]["min"].view(np.uint32) == n["min_bits"]).all(), (c["name"], k)
            assert (t["nodes"][k]["max"].view(np.uint32) == n["max_bits"]).all(), (c["name"], k)
            assert (t["nodes"][k]["left"], t["nodes"][k]["skip"]) == (n["left"], n["skip"])
        assert (t["store"] == np.array(c["store_serial"])).all(), c["name"]
        # the own-leaf form (TreeData) overwrites only the root box
        _, t2 = O.build_traverse_bvh(x, c["apl"], c["r"], mode=O.QUERY_OWN_LEAF)
        nL = len(x) // c["apl"]
        root = t2["nodes"][nL]
        assert (root["min"] == 0).all() and (root["max"] == 1).all()
        same = np.ones(len(t2["nodes"]), bool)
        same[nL] = False
        assert (t2["nodes"][same] == t["nodes"][same]).all()


def test_ranks_match_golden(O, pq_golden):
    for c in pq_golden:
        assert (O.quantized_positions(_x(c)) == np.array(c["quantized"])).all(), c["name"]


def test_position8_every_atom_paired(O, pq_golden):
    """test/BVHTraverse.jl:181-198: at r = 10 every one of the 8 atoms is
    paired, and the list equals the all-pairs list as a set."""
    c = pq_golden[0]
    assert c["name"] == "position8_apl1_r10"
    p, _ = O.build_traverse_bvh(_x(c), 1, 10.0)
    assert len(p) == 28
    assert set(p["i"].tolist()) | set(p["j"].tolist()) == set(range(8))
    assert _unordered(p) == {(i, j) for i in range(8) for j in range(i + 1, 8)}


def test_margin_pairs_fall_outside_the_box(O, pq_golden):
    """The two margin pairs pass s < fl(r*r) but sit on their box boundary,
    so the strict point-in-box test drops them (a non-strict test keeps them)."""
    c = next(c for c in pq_golden if c["name"].startswith("margin"))
    x = _x(c)
    b = O.bruteforce(x, c["r"])
    p, _ = O.build_traverse_bvh(x, 1, c["r"])
    assert {(0, 1), (2, 3)} <= _unordered(b)
    assert not ({(0, 1), (2, 3)} & _unordered(p))
    assert _unordered(p) == _unordered(b) - {(0, 1), (2, 3)}


def test_query_pairs_keep_the_lower_sorted_atom_first(O):
    """Own-leaf form: every pair is (query, subject) with the query earlier
    in Morton order; root form: the lower original index first."""
    x = np.random.default_rng(5).random((400, 3)).astype(np.float32)
    p, t = O.build_traverse_bvh(x, 4, 0.15, mode=O.QUERY_OWN_LEAF)
    rank = np.empty(400, np.int64)
    rank[t["perm"]] = np.arange(400)
    assert (rank[p["i"]] < rank[p["j"]]).all()
    q, _ = O.build_traverse_bvh(x, 4, 0.15, mode=O.QUERY_ROOT)
    assert (q["i"] < q["j"]).all()
    # query-major order: queries in increasing sorted position
    assert (np.diff(rank[p["i"]]) >= 0).all()
    assert (np.diff(rank[q["i"]]) >= 0).all()


@pytest.mark.parametrize("r", [1e-4, 1e-3, 1e-2, 1e-1, 1.0])
def test_equals_all_pairs_radius_sweep(O, r):
    """test/BVHTraverse.jl:203-222 (radius 1e-4 .. 1 over 5,000 points)."""
    x = np.random.default_rng(21).random((5000, 3)).astype(np.float32)
    p, _ = O.build_traverse_bvh(x, 1, r, nthreads=4)
    assert _same_pairs(p, O.bruteforce(x, r))


@pytest.mark.parametrize("apl", [1, 2, 4, 5, 10, 20, 1000])
def test_equals_all_pairs_leaf_size_sweep(O, apl):
    """test/BVHTraverse.jl:224-243 (leaf sizes 1 .. 1000, r = 0.1)."""
    x = np.random.default_rng(22).random((5000, 3)).astype(np.float32)
    for mode in (O.QUERY_OWN_LEAF, O.QUERY_ROOT):
        p, _ = O.build_traverse_bvh(x, apl, 0.1, mode=mode, nthreads=4)
        assert _same_pairs(p, O.bruteforce(x, 0.1))


@pytest.mark.parametrize("n", [10, 100, 1000, 2000, 3000, 20000])
def test_equals_all_pairs_size_sweep(O, n):
    """test/BVHTraverse.jl:245-264 (10 .. 20,000 points, apl 5, r = 0.1)."""
    x = np.random.default_rng(23 + n).random((n, 3)).astype(np.float32)
    p, _ = O.build_traverse_bvh(x, 5, 0.1, nthreads=4)
    if n <= 3000:
        assert _same_pairs(p, O.bruteforce(x, 0.1))
    else:  # cell-binned all-pairs check
        assert _unordered(p) == _cell_pairs(x, 0.1)


def _cell_pairs(x, r):
    r2 = np.float32(r) * np.float32(r)
    cells = np.floor(x / r).astype(np.int64)
    key = {}
    for i, c in enumerate(map(tuple, cells)):
        key.setdefault(c, []).append(i)
    out = set()
    for c, members in key.items():
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    other = key.get((c[0] + dx, c[1] + dy, c[2] + dz))
                    if not other:
                        continue
                    for i in members:
                        d = x[other] - x[i]
                        s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
                        for j in np.asarray(other)[s < r2]:
                            if i < j:
                                out.add((i, int(j)))
    return out


def test_threaded_order_is_a_permutation(O):
    x = np.random.default_rng(8).random((3000, 3)).astype(np.float32)
    one, _ = O.build_traverse_bvh(x, 4, 0.1, nthreads=1)
    many, _ = O.build_traverse_bvh(x, 4, 0.1, nthreads=7)
    assert (O.canonical(one) == O.canonical(many)).all()


def test_threaded_store_holds_a_child_range_end(O):
    """With several threads a slot holds whichever child arrived first: the
    left child's first leaf or the right child's last leaf."""
    x = np.random.default_rng(9).random((4096, 3)).astype(np.float32)
    _, t1 = O.build_traverse_bvh(x, 4, 0.05, nthreads=1)
    _, t8 = O.build_traverse_bvh(x, 4, 0.05, nthreads=8)
    ranges = node_ranges(t1["nodes"], 1024)
    check_store(t8["store"], ranges)
    assert (t1["store"] == np.array([ranges[k][0] for k in range(1, 1024)])).all()


def node_ranges(nodes, nL):
    """split k (1-based) -> (first leaf, last leaf) of the internal node that
    splits after leaf k, from the tree's left/skip links."""
    first = {}
    last = {}

    def walk(k):
        if k <= nL:
            return k, k
        left = nodes[k - 1]["left"]
        right = nodes[left - 1]["skip"]
        a, b = walk(left)
        c, d = walk(right)
        first[b] = (a, d)
        return a, d

    import sys
    sys.setrecursionlimit(10000)
    walk(nL + 1)
    return first


def check_store(store, ranges):
    for k in range(1, len(store) + 1):
        assert store[k - 1] in ranges[k], (k, store[k - 1], ranges[k])


def test_ranks_invariants(O):
    rng = np.random.default_rng(10)
    x = rng.choice(np.array([-1.0, -0.0, 0.0, 0.5, 2.0], np.float32), size=(500, 3))
    x[::7, 1] = rng.random(len(x[::7])).astype(np.float32)
    q = O.quantized_positions(x)
    for d in range(3):
        assert sorted(q[:, d]) == list(range(1, 501))
        o = np.argsort(q[:, d])
        v = x[o, d]
        key = np.where(v == 0, np.where(np.signbit(v), -1e-30, 1e-30), v)
        assert (np.diff(key) >= 0).all()  # monotone, -0 before +0
    # ties on x keep the identity order; ties on y keep the x order
    ox = np.argsort(q[:, 0])
    for a, b in zip(ox[:-1], ox[1:]):
        if x[a, 0] == x[b, 0] and np.signbit(x[a, 0]) == np.signbit(x[b, 0]):
            assert a < b
    oy = np.argsort(q[:, 1])
    for a, b in zip(oy[:-1], oy[1:]):
        if x[a, 1] == x[b, 1] and np.signbit(x[a, 1]) == np.signbit(x[b, 1]):
            assert q[a, 0] < q[b, 0]
