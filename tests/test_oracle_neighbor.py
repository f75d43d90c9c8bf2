"""CPU: pins the neighbour-search oracle (oracle/refcpu.c).

Checks against the golden fixtures (an independent top-down restatement,
tests/golden/make_golden.py), hand-derived keys of the reference test's position8 points,
numpy's stable sort, brute force (AllToAll.jl semantics, as the reference's
own tests do: test/BVHTraverse.jl:205-209, 222-289) and tree invariants.
"""
import numpy as np
import pytest
from conftest import bruteforce_np, pairs_to_rows, rows_canonical

POSITION8 = np.array([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.346, 0.98, 0.12],
                      [0.01, 0.76, 0.99], [0.1111, 0.4, 0.31], [0.234, 0.29, 0.2],
                      [0.11346, 0.918, 0.1276], [0.061, 0.76, 0.989]], np.float32)


def test_position8_keys_known_answer(O):
    # Derived by hand from BVHTraverse.jl:184,241,266-269 (q = floor(x/binwidth)).
    q = np.array([[102, 102, 102], [204, 204, 204], [353, 1002, 122], [10, 777, 1012],
                  [113, 409, 317], [239, 296, 204], [116, 939, 130], [62, 777, 1011]])
    assert np.array_equal(np.floor(POSITION8 / O.binwidth()).astype(int), q)
    assert O.morton_masked(POSITION8).tolist() == [102, 204, 227, 300, 501, 77, 194, 296]
    # Julia sortperm (1-based) [6,1,7,2,3,8,4,5]
    assert (O.sortperm(O.morton_masked(POSITION8)) + 1).tolist() == [6, 1, 7, 2, 3, 8, 4, 5]


def test_binwidth_is_rounded_once(O):
    # Float32(1/1023): quotient in Float64, one rounding (BVHTraverse.jl:184)
    assert O.binwidth() == np.float32(1.0 / 1023.0)
    assert O.binwidth().view(np.uint32) == 0x3A802008


def test_masked_key_bit_invariants(O):
    rng = np.random.default_rng(5)
    x = (rng.random((20000, 3)) * 300.0 - 50.0).astype(np.float32)
    c = O.morton_masked(x).astype(np.int64)
    q = np.floor(x / O.binwidth()).astype(np.int64)
    for d, m in enumerate([0x09249249, 0x12492492, 0x24924924]):
        assert np.array_equal(c & m, q[:, d] & m)
    assert np.all(c >= 0) and np.all(c < 2 ** 30)  # masks never set bit 30/31


def test_masked_key_domain_error(O):
    for bad in [np.nan, np.inf, -np.inf, 3.0e6, -3.0e6]:
        x = np.zeros((4, 3), np.float32)
        x[2, 1] = bad
        with pytest.raises(ValueError):
            O.morton_masked(x)
    # largest accepted: floor(v) = 2^31 - 128 (v < 2^31)
    x = np.zeros((1, 3), np.float32)
    x[0, 0] = np.float32(2147483520.0) * O.binwidth()
    O.morton_masked(x)


def test_interleave_key_is_bit_interleave(O):
    rng = np.random.default_rng(6)
    x = (rng.random((5000, 3)) * 40.0).astype(np.float32)
    lo, sc = O.interleave_scale([0, 0, 0], [40, 40, 40])
    c = O.morton_interleave(x, lo, sc).astype(np.int64)
    t = (x - lo) * sc
    q = np.clip(np.floor(t), 0, 1023).astype(np.int64)
    for d in range(3):
        dec = np.zeros_like(c)
        for k in range(10):
            dec |= ((c >> (3 * k + d)) & 1) << k
        assert np.array_equal(dec, q[:, d])


def test_sortperm_stable_matches_numpy(O):
    rng = np.random.default_rng(7)
    codes = rng.integers(-50, 50, size=10000).astype(np.int32)  # many ties, signed
    assert np.array_equal(O.sortperm(codes), np.argsort(codes, kind="stable"))
    codes = rng.integers(-2 ** 31, 2 ** 31 - 1, size=10000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(O.sortperm(codes), np.argsort(codes, kind="stable"))


def _tree_from_case(O, c, nthreads=1):
    x = c["x"]
    codes = O.morton_masked(x)
    perm = O.sortperm(codes)
    return O.build_tree(x[perm], codes[perm], c["apl"], c["r"], True, nthreads), perm


@pytest.mark.parametrize("nthreads", [1, 8])
def test_tree_matches_topdown_golden(O, golden, nthreads):
    for c in golden:
        nodes, perm = _tree_from_case(O, c, nthreads)
        assert perm.tolist() == c["perm"], c["name"]
        assert O.morton_masked(c["x"]).tolist() == c["codes"], c["name"]
        g = c["nodes"]
        assert len(nodes) == len(g)
        assert nodes["left"].tolist() == [n["left"] for n in g], c["name"]
        assert nodes["skip"].tolist() == [n["skip"] for n in g], c["name"]
        assert np.array_equal(nodes["min"].view(np.uint32), np.array([n["min_bits"] for n in g], np.uint32))
        assert np.array_equal(nodes["max"].view(np.uint32), np.array([n["max_bits"] for n in g], np.uint32))


def test_position8_rope_pattern(O, golden):
    # test/BVHTraverse.jl:184-187: keys[8].skip == 0, keys[1..7].skip != 0
    c = golden[0]
    nodes, _ = _tree_from_case(O, c)
    assert nodes["skip"][7] == 0 and np.all(nodes["skip"][:7] != 0)


def _check_traversable(nodes, nL):
    # every node reachable exactly once from the root by left/skip in DFS order
    seen = np.zeros(2 * nL, bool)
    t = nodes["left"][nL]  # root's left (root itself is never visited by leaf queries)
    seen[nL + 1] = True
    order = []
    while t != 0:
        assert not seen[t], "revisit"
        seen[t] = True
        n = nodes[t - 1]
        if n["left"] == 0:
            order.append(t)
            t = n["skip"]
        else:
            t = n["left"]
    assert seen[1:].all()
    assert order == list(range(1, nL + 1))  # leaves in index order


def test_tree_invariants_random(O):
    rng = np.random.default_rng(8)
    for N, apl, L in [(4096, 4, 17.2), (3000, 5, 1.0), (1000, 1, 1.0), (6000, 20, 5.0)]:
        x = (rng.random((N, 3)) * L).astype(np.float32)
        codes = O.morton_masked(x)
        perm = O.sortperm(codes)
        nodes = O.build_tree(x[perm], codes[perm], apl, 0.3, False, 8)
        nL = N // apl
        _check_traversable(nodes, nL)
        # parent box == exact union of its two children (bitwise)
        for k in range(nL + 1, 2 * nL):
            n = nodes[k - 1]
            lc = n["left"]
            # right child = the node reached by the left child's rope
            rc = nodes[lc - 1]["skip"]
            for f, op in (("min", np.minimum), ("max", np.maximum)):
                assert np.array_equal(n[f], op(nodes[lc - 1][f], nodes[rc - 1][f]))
        # determinism across thread counts
        nodes1 = O.build_tree(x[perm], codes[perm], apl, 0.3, False, 1)
        assert nodes.tobytes() == nodes1.tobytes()


def test_golden_pairs_in_single_thread_order(O, golden):
    for c in golden:
        for pred, key in ((O.PRED_SCALAR, "pairs_scalar"), (O.PRED_SIMD, "pairs_simd")):
            p, _, _ = O.leafbuild_traverse_bvh(c["x"], c["apl"], c["r"], predicate=pred, nthreads=1)
            assert np.array_equal(pairs_to_rows(p), c[key]), (c["name"], key)
        b = O.bruteforce(c["x"], c["r"])
        assert np.array_equal(pairs_to_rows(b), c["bruteforce"]), c["name"]


def test_threaded_traversal_same_set(O):
    rng = np.random.default_rng(9)
    x = (rng.random((5000, 3))).astype(np.float32)
    p1, _, _ = O.leafbuild_traverse_bvh(x, 4, 0.1, nthreads=1)
    p8, _, _ = O.leafbuild_traverse_bvh(x, 4, 0.1, nthreads=8)
    assert np.array_equal(O.canonical(p1), O.canonical(p8))


@pytest.mark.parametrize("r", [1e-4, 1e-3, 1e-2, 1e-1, 1.0])
def test_bvh_equals_bruteforce_radius_sweep(O, r):
    # test/BVHTraverse.jl:225-243 (pos5000, radius sweep)
    x = np.random.default_rng(10).random((5000, 3)).astype(np.float32)
    b = O.bruteforce(x, r)
    p, _, _ = O.leafbuild_traverse_bvh(x, 5, r, predicate=O.PRED_SIMD, nthreads=8)
    assert np.array_equal(O.canonical(p), O.canonical(b))


@pytest.mark.parametrize("apl", [1, 2, 4, 5, 10, 20, 1000])
def test_bvh_equals_bruteforce_apl_sweep(O, apl):
    # test/BVHTraverse.jl:246-265 (apl sweep at r = 0.1)
    x = np.random.default_rng(11).random((5000, 3)).astype(np.float32)
    b = O.bruteforce(x, 0.1)
    p, _, _ = O.leafbuild_traverse_bvh(x, apl, 0.1, predicate=O.PRED_SIMD, nthreads=8)
    assert np.array_equal(O.canonical(p), O.canonical(b))


@pytest.mark.parametrize("N", [10, 100, 1000, 2000, 3000, 20000])
def test_bvh_equals_bruteforce_size_sweep(O, N):
    # test/BVHTraverse.jl:268-287 (sizes at apl = 5, r = 0.1)
    x = np.random.default_rng(12 + N).random((N, 3)).astype(np.float32)
    b = O.bruteforce(x, 0.1)
    p, _, _ = O.leafbuild_traverse_bvh(x, 5, 0.1, predicate=O.PRED_SIMD, nthreads=8)
    assert np.array_equal(O.canonical(p), O.canonical(b))


def margin_points(count=8, r=0.1, seed=0):
    """Point pairs at distance ~r whose float32 squared distance s satisfies
    s < fl(r*r) while fl(sqrt(s)) == r (or the reverse), so the scalar
    (BVHTraverse.jl:1027) and SIMD (NaiveSIMD.jl:589) predicates disagree."""
    F = np.float32
    r = F(r)
    r2 = F(r * r)
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        a = rng.random(3).astype(F) * F(0.5) + F(0.2)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        b = (a.astype(np.float64) + u * float(r) * (1 + rng.uniform(-3e-7, 3e-7))).astype(F)
        d = (a - b).astype(F)
        s = F(F(F(d[0] * d[0]) + F(d[1] * d[1])) + F(d[2] * d[2]))
        if (s < r2) != (F(np.sqrt(s)) < r):
            out.append((a, b))
    return r, out


def margin_set(r=0.1, count=8, seed=0):
    """Margin pairs spread apart (offset by 1.0 each) plus filler atoms."""
    r, cases = margin_points(count, r, seed)
    pts = []
    for k, (a, b) in enumerate(cases):
        off = np.float32(1.0 * k)
        pts.append(a + off)
        pts.append(b + off)
    rng = np.random.default_rng(seed + 1)
    filler = (rng.random((64 - len(pts), 3)) * 8.0).astype(np.float32)
    return r, np.concatenate([np.array(pts, np.float32), filler]).astype(np.float32)


def test_margin_predicates_differ(O):
    r, pts = margin_set()
    ps, _, _ = O.leafbuild_traverse_bvh(pts, 2, r, predicate=O.PRED_SCALAR)
    pv, _, _ = O.leafbuild_traverse_bvh(pts, 2, r, predicate=O.PRED_SIMD)
    b = O.bruteforce(pts, r)
    assert len(ps) != len(pv)  # the predicates disagree on the margin pairs
    assert np.array_equal(O.canonical(pv), O.canonical(b))  # brute force is sqrt(s) < r too


def test_canonical_rows_helper():
    a = np.array([[3, 1, 7], [0, 2, 5]])
    assert rows_canonical(a).tolist() == [[0, 2, 5], [1, 3, 7]]


# ---- Hilbert key (the simulation's key): pinned by the curve's defining
# properties, not by a restatement of the transform ----

def _grid(k):
    """All cells of a 2^k grid, scaled to 10-bit cell coordinates."""
    g = np.arange(2 ** k)
    c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return c << (10 - k)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_hilbert_full_grid_is_a_face_adjacent_path(O, k):
    c = _grid(k)
    h = O.hilbert_cells(c).astype(np.int64)
    # on a 2^k grid the top 3k bits of the index run through 0 .. 8^k - 1 once
    top = h >> (3 * (10 - k))
    assert np.array_equal(np.sort(top), np.arange(8 ** k))
    # consecutive cells along the curve share a face (one unit step)
    path = c[np.argsort(h)] >> (10 - k)
    steps = np.abs(np.diff(path, axis=0)).sum(1)
    assert np.all(steps == 1)


def test_hilbert_prefixes_are_octree_cubes(O):
    # cells with the same 3k-bit prefix lie in one cube of side 2^(10-k)
    rng = np.random.default_rng(8)
    c = rng.integers(0, 1024, (50000, 3))
    h = O.hilbert_cells(c).astype(np.int64)
    for k in (1, 3, 5):
        pre = h >> (3 * (10 - k))
        cube = c >> (10 - k)
        for p in np.unique(pre)[:200]:
            m = pre == p
            assert (cube[m] == cube[m][0]).all()


def test_hilbert_is_a_bijection_and_starts_at_the_origin(O):
    rng = np.random.default_rng(9)
    c = np.unique(rng.integers(0, 1024, (200000, 3)), axis=0)
    h = O.hilbert_cells(c)
    assert len(np.unique(h)) == len(c)
    assert O.hilbert_cells(np.zeros((1, 3), np.int64))[0] == 0
    assert np.all((h >= 0) & (h < 2 ** 30))


def test_morton_hilbert_uses_the_interleaved_cells(O):
    rng = np.random.default_rng(10)
    x = (rng.random((30000, 3)) * 68.94).astype(np.float32)
    lo, sc = O.interleave_scale([0, 0, 0], [68.94, 68.94, 68.94])
    t = (x - lo) * sc
    cells = np.where(t <= 0, 0, np.where(t >= 1023, 1023, np.floor(t))).astype(np.int64)
    assert np.array_equal(O.morton_hilbert(x, lo, sc), O.hilbert_cells(cells))


def test_numpy_bruteforce_helper_matches_oracle(O):
    """The tests' definition-level brute force (both predicates) against the
    oracle's brute force (sqrt(s) < r, AllToAll.jl:65-88) and its scalar-test
    traversal on the margin set, where the two predicates disagree."""
    r, pts = margin_set()
    simd = bruteforce_np(pts, r, 1)
    assert np.array_equal(simd, pairs_to_rows(O.bruteforce(pts, r)))
    ps, _, _ = O.leafbuild_traverse_bvh(pts, 2, r, predicate=O.PRED_SCALAR)
    assert np.array_equal(bruteforce_np(pts, r, 0), rows_canonical(pairs_to_rows(ps)))
    assert len(bruteforce_np(pts, r, 0)) != len(simd)
    x = (np.random.default_rng(3).random((700, 3)) * 6.0).astype(np.float32)
    assert np.array_equal(bruteforce_np(x, 1.0, 1), pairs_to_rows(O.bruteforce(x, 1.0)))
