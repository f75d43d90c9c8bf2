"""GPU: the point-query traversal (TreeData + neighbor_traverse,
BVHTraverse.jl:500-543, 1125-1234; the root-started walk of
gpuneighbor_traverse!, ext/NaiveKA.jl:352-452), TreeData's store,
quantized_positions! (BVHTraverse.jl:290-314) and the full treedata of
gpubvh_neighborlist (ext/NaiveKA.jl:557).

Bar: pairs bit-exact and in the oracle's single-thread order (index, index,
distance bits); trees bit-exact; ranks exact.  The store is filled by racing
lanes, so each slot is checked to hold one of its two legal values (the left
child's first leaf or the right child's last leaf, read off the tree).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from test_oracle_point_query import check_store, node_ranges

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = {"uniform": 101, "sizes": 102, "stage": 103, "ranks": 104, "big": 105}


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


def _spec(pkg, n, apl, r):
    return pkg.SpheresBVHSpecs(neighbor_distance=r, atom_count=n, atomsperleaf=apl)


@pytest.mark.parametrize("query,key", [(0, "pairs_own_leaf"), (1, "pairs_root")])
def test_golden_pairs_trees_ranks(pkg, pq_golden, query, key):
    for c in pq_golden:
        x = _x(c)
        n, apl = len(x), c["apl"]
        spec = _spec(pkg, n, apl, c["r"])
        print("case", c["name"], "query", query, flush=True)
        p = pkg.bvh.build_traverse_bvh(x, spec, query=query)
        g = np.array(c[key], np.int64).reshape(-1, 3)
        assert _rows(p).shape == g.shape and (_rows(p) == g).all(), c["name"]
        nodes, prim, perm = pkg.bvh.tree_export(spec, x)
        assert (perm == np.array(c["perm"])).all()
        nL = n // apl
        for k, gn in enumerate(c["nodes_root_kept"]):
            if query == 0 and k == nL:  # TreeData overwrites the root box
                assert (nodes[k]["min"] == 0).all() and (nodes[k]["max"] == 1).all()
            else:
                assert (nodes[k]["min"].view(np.uint32) == gn["min_bits"]).all(), (c["name"], k)
                assert (nodes[k]["max"].view(np.uint32) == gn["max_bits"]).all(), (c["name"], k)
            assert (nodes[k]["left"], nodes[k]["skip"]) == (gn["left"], gn["skip"])
        check_store(pkg.bvh.tree_store(spec), node_ranges(nodes, nL))
        assert (pkg.bvh.quantized_positions(x) == np.array(c["quantized"])).all()


@pytest.mark.parametrize("apl", [1, 2, 4, 5, 16, 64])
@pytest.mark.parametrize("query", [0, 1])
def test_pairs_match_oracle_in_order(pkg, O, apl, query):
    n = 20480
    x = np.random.default_rng(SEEDS["uniform"] + apl).random((n, 3)).astype(np.float32)
    print("uniform", n, "apl", apl, "query", query, flush=True)
    p = pkg.bvh.build_traverse_bvh(x, _spec(pkg, n, apl, 0.08), query=query)
    ref, tree = O.build_traverse_bvh(x, apl, 0.08, mode=query)
    assert len(p) == len(ref) > 0
    assert (_rows(p) == _rows(ref)).all()
    nodes, _, perm = pkg.bvh.tree_export(_spec(pkg, n, apl, 0.08), x)
    assert (perm == tree["perm"]).all()
    assert (nodes.view(np.uint8) == tree["nodes"].view(np.uint8)).all()


@pytest.mark.parametrize("n,apl,r", [(8, 1, 10.0), (10, 5, 0.1), (1000, 5, 0.1), (3000, 4, 1e-3),
                                     (5000, 1000, 0.1), (65536, 8, 0.03)])
def test_sizes_and_edges(pkg, O, n, apl, r):
    """Leaf sizes, radii and sizes of test/BVHTraverse.jl:203-280, plus two
    leaves of 1,000 atoms and a large case."""
    x = np.random.default_rng(SEEDS["sizes"] + n).random((n, 3)).astype(np.float32)
    print("sizes", n, apl, r, flush=True)
    for query in (0, 1):
        p = pkg.bvh.build_traverse_bvh(x, _spec(pkg, n, apl, r), query=query)
        ref, _ = O.build_traverse_bvh(x, apl, r, mode=query)
        assert (_rows(p) == _rows(ref)).all(), query


def test_stage_traversal_over_oracle_tree(pkg, O):
    """neighbor_traverse as a stage: the oracle's TreeData tree, sorted
    positions and indices in, the oracle's walk out."""
    n, apl, r = 12000, 4, 0.09
    x = np.random.default_rng(SEEDS["stage"]).random((n, 3)).astype(np.float32)
    ref, tree = O.build_traverse_bvh(x, apl, r)
    prim = {"index": tree["perm"], "position": x[tree["perm"]]}
    for query in (0, 1):
        p = pkg.bvh.neighbor_traverse(tree["nodes"], prim, _spec(pkg, n, apl, r), query=query)
        want = O.neighbor_traverse(tree["nodes"], x[tree["perm"]], tree["perm"], apl, r, mode=query)
        assert (_rows(p) == _rows(want)).all(), query
    assert (_rows(p) != 0).any()


def test_treedata_and_gpubvh_neighborlist(pkg, O):
    n, apl, r = 16384, 4, 0.07
    x = np.random.default_rng(SEEDS["stage"] + 1).random((n, 3)).astype(np.float32)
    spec = _spec(pkg, n, apl, r)
    keys, prim, mortsort, store = pkg.TreeData(x, spec)
    _, tree = O.build_traverse_bvh(x, apl, r)
    assert (keys.view(np.uint8) == tree["nodes"].view(np.uint8)).all()
    assert (mortsort == tree["perm"]).all() and (prim["position"] == x[mortsort]).all()
    check_store(store, node_ranges(keys, n // apl))
    res = pkg.gpubvh_neighborlist("hip", x, spec)
    td = res["treedata"]
    assert set(td) == {"tree", "mortonpositions", "quantizedposition", "store"}
    assert all(v is not None for v in td.values())
    ref, rtree = O.build_traverse_bvh(x, apl, r, mode=O.QUERY_ROOT)
    assert (_rows(res["pairlist"]) == _rows(ref)).all()
    assert (td["tree"].view(np.uint8) == rtree["nodes"].view(np.uint8)).all()
    assert (td["quantizedposition"] == O.quantized_positions(x)).all()
    check_store(td["store"], node_ranges(td["tree"], n // apl))


def test_ranks_match_oracle_with_ties(pkg, O):
    rng = np.random.default_rng(SEEDS["ranks"])
    n = 1 << 20
    x = rng.random((n, 3)).astype(np.float32)
    # coarse coordinates on y (many ties) and signed zeros on z
    x[:, 1] = np.floor(x[:, 1] * 64) / 64
    x[: n // 4, 2] = rng.choice(np.array([-0.0, 0.0], np.float32), n // 4)
    q = pkg.bvh.quantized_positions(x)
    assert (q == O.quantized_positions(x)).all()


def test_capacity_protocol(pkg):
    n, apl, r = 4096, 4, 0.1
    x = np.random.default_rng(SEEDS["big"]).random((n, 3)).astype(np.float32)
    ctx = pkg.default_context()
    d = pkg.DeviceArray.from_host(ctx, x)
    spec = _spec(pkg, n, apl, r)
    lib = pkg.library()
    count = ctypes.c_int64(0)
    out = pkg.DeviceArray(ctx, (4,), pkg.PAIR_DTYPE)
    rc = lib.nd_build_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), 0, 0, out.ptr, 4,
                                   ctypes.byref(count))
    assert rc == pkg._lib.ND_ERR_CAPACITY and count.value > 4
    full = pkg.bvh.build_traverse_bvh(x, spec)
    assert len(full) == count.value
    rc = lib.nd_build_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), 7, 0, out.ptr, 4,
                                   ctypes.byref(count))
    assert rc == pkg._lib.ND_ERR_ARG


def test_full_size_count_matches_cell_list(pkg):
    """262,144 atoms of the uniform liquid (density 0.8) at r = 2.8: the point
    query keeps exactly the cell list's pairs with s < fl(r*r) (up to pairs on
    a box boundary, none in random data)."""
    n = 262144
    L = pkg.synth.box_side(n)
    x = pkg.synth.lattice_positions(n, L, SEEDS["big"]).astype(np.float32) / np.float32(L)
    r = np.float32(2.8 / L)
    spec = _spec(pkg, n, 16, r)
    p = pkg.bvh.build_traverse_bvh(x, spec)
    ctx = pkg.default_context()
    count = pkg.cells.neighborlist_device(pkg.DeviceArray.from_host(ctx, x), r, pkg.PRED_SCALAR)[1]
    assert len(p) == count
    assert (p["i"] != p["j"]).all() and (p["d"] < r).all()
