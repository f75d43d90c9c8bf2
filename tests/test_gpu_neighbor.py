"""GPU parity of the neighbour-search path against the oracle: keys, sort
order, tree nodes and pair lists bit-exact (including emission order, which
equals a single-threaded reference run)."""
import numpy as np
import pytest
from conftest import pairs_to_rows, rows_canonical
from test_oracle_neighbor import margin_set

pytestmark = pytest.mark.gpu


def _rand(n, L, seed, lo=0.0):
    return (np.random.default_rng(seed).random((n, 3)) * L + lo).astype(np.float32)


def test_masked_keys_bitexact(pkg, O):
    for seed, (L, lo) in enumerate([(1.0, 0.0), (68.94, 0.0), (500.0, -250.0), (4000.0, -100.0)]):
        x = _rand(50000, L, seed, lo)
        assert np.array_equal(pkg.mortoncodes(x), O.morton_masked(x)), (L, lo)
    # every quantisation cell edge for small q: x = k*binwidth and its neighbours
    k = np.arange(-3000, 3000, dtype=np.float64)
    bw = np.float32(O.binwidth())
    e = (k * bw).astype(np.float32)
    cand = np.concatenate([e, np.nextafter(e, np.float32(np.inf)), np.nextafter(e, np.float32(-np.inf))])
    x = np.stack([cand, cand[::-1], np.roll(cand, 7)], axis=1).astype(np.float32)
    assert np.array_equal(pkg.mortoncodes(x), O.morton_masked(x))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf, 3.0e6, -3.0e6])
def test_masked_keys_domain_error(pkg, bad):
    x = _rand(1000, 1.0, 3)
    x[517, 2] = bad
    with pytest.raises(pkg.DomainError):
        pkg.mortoncodes(x)
    pkg.mortoncodes(_rand(1000, 1.0, 3))  # the context is usable afterwards


def test_interleave_keys_bitexact(pkg, O):
    x = _rand(70000, 68.94, 4)
    lo, sc = O.interleave_scale([0, 0, 0], [68.94, 68.94, 68.94])
    got = pkg.mortoncodes(x, mode=pkg.MORTON_INTERLEAVE, lo=lo, scale=sc)
    assert np.array_equal(got, O.morton_interleave(x, lo, sc))


def test_hilbert_keys_bitexact(pkg, O):
    x = _rand(70000, 68.94, 41)
    lo, sc = O.interleave_scale([0, 0, 0], [68.94, 68.94, 68.94])
    got = pkg.mortoncodes(x, mode=pkg.MORTON_HILBERT, lo=lo, scale=sc)
    assert np.array_equal(got, O.morton_hilbert(x, lo, sc))
    # box from the data (no lo/scale): same rule as the interleaved key
    y = _rand(5000, 3.0, 42)
    ylo = y.min(axis=0)
    ysc = (np.float32(1024.0) / (y.max(axis=0) - ylo)).astype(np.float32)
    assert np.array_equal(pkg.mortoncodes(y, mode=pkg.MORTON_HILBERT), O.morton_hilbert(y, ylo, ysc))


@pytest.mark.parametrize("n", [1, 2, 1000, 1024, 1025, 262144, 1000003])
def test_sortperm_stable(pkg, O, n):
    rng = np.random.default_rng(n)
    codes = rng.integers(-(2 ** 31), 2 ** 31 - 1, size=n, dtype=np.int64).astype(np.int32)
    codes[: n // 2] = codes[: n // 2] % 97  # heavy ties
    cs, perm = pkg.sortperm(codes)
    ref = O.sortperm(codes)
    assert np.array_equal(perm, ref)
    assert np.array_equal(cs, codes[ref])


def test_tree_golden_bitexact(pkg, golden):
    for c in golden:
        x = c["x"]
        perm = np.array(c["perm"])
        codes = np.array(c["codes"], np.int32)
        spec = pkg.SpheresBVHSpecs(neighbor_distance=c["r"], atom_count=len(x), atomsperleaf=c["apl"])
        nodes = pkg.bounding_volume_hierarchy(x[perm], codes[perm], spec)
        g = c["nodes"]
        assert nodes["left"].tolist() == [n["left"] for n in g], c["name"]
        assert nodes["skip"].tolist() == [n["skip"] for n in g], c["name"]
        assert np.array_equal(nodes["min"].view(np.uint32), np.array([n["min_bits"] for n in g], np.uint32))
        assert np.array_equal(nodes["max"].view(np.uint32), np.array([n["max_bits"] for n in g], np.uint32))


def test_golden_pairs_exact_order(pkg, golden):
    for c in golden:
        spec = pkg.SpheresBVHSpecs(neighbor_distance=c["r"], atom_count=len(c["x"]),
                                   atomsperleaf=c["apl"])
        got = pkg.leafbuild_traverse_bvh(c["x"], spec)
        assert np.array_equal(pairs_to_rows(got), c["pairs_scalar"]), c["name"]
        if c["apl"] == 4:
            got = pkg.simdbuild_traverse_bvh(c["x"], spec)
            assert np.array_equal(pairs_to_rows(got), c["pairs_simd"]), c["name"]
        b = pkg.threshold_unique_pairs(c["x"], c["r"])
        assert np.array_equal(pairs_to_rows(b), c["bruteforce"]), c["name"]


def _compare_with_oracle(pkg, O, x, apl, r, pred, morton=0, lo=None, scale=None):
    spec = pkg.SpheresBVHSpecs(neighbor_distance=r, atom_count=len(x), atomsperleaf=apl)
    out, n = pkg.neighborlist_device(x, spec, pred, morton)
    got = out.to_host(n)
    ref, tree, _ = O.leafbuild_traverse_bvh(x, apl, r, predicate=pred, morton=morton, lo=lo,
                                            scale=scale, want_tree=True)
    assert len(got) == len(ref)
    assert np.array_equal(pairs_to_rows(got), pairs_to_rows(ref))
    nodes, prim, perm = pkg.bvh.tree_export(spec, x)
    assert np.array_equal(perm, tree["perm"])
    assert nodes.tobytes() == tree["nodes"].tobytes()
    return got


@pytest.mark.parametrize("r", [1e-4, 1e-3, 1e-2, 1e-1, 1.0])
def test_radius_sweep_vs_oracle(pkg, O, r):
    x = _rand(5000, 1.0, 10)
    _compare_with_oracle(pkg, O, x, 5, r, pkg.PRED_SCALAR)


@pytest.mark.parametrize("apl", [1, 2, 4, 5, 10, 20, 1000])
def test_apl_sweep_vs_oracle(pkg, O, apl):
    x = _rand(5000, 1.0, 11)
    _compare_with_oracle(pkg, O, x, apl, 0.1, pkg.PRED_SCALAR)
    if apl == 4:
        _compare_with_oracle(pkg, O, x, apl, 0.1, pkg.PRED_SIMD)


@pytest.mark.parametrize("apl", [8, 16, 32, 64])
def test_wave_walk_leaf_sizes_vs_oracle(pkg, O, apl):
    """Leaf sizes dividing 64 take the one-walk-per-wave traversal (64 / apl
    query leaves per wave); the pairs and their order are the rope walk's."""
    L = (4096 / 0.8) ** (1 / 3)
    x = _rand(4096, L, 19 + apl)
    for pred in (pkg.PRED_SCALAR, pkg.PRED_SIMD):
        _compare_with_oracle(pkg, O, x, apl, 1.5, pred)


# fixed seeds, one per input shape (reproducible in every process)
SHAPE_SEEDS = {"duplicates": 101, "line": 102, "plane": 103, "two_clusters": 104}


@pytest.mark.parametrize("shape", ["duplicates", "line", "plane", "two_clusters"])
@pytest.mark.parametrize("morton", [0, 1])
def test_degenerate_inputs_vs_oracle(pkg, O, shape, morton):
    """Collisions (64 distinct points, each repeated 64 times: equal keys, ties
    broken by index, pairs at distance 0), points on a line or a plane (keys
    varying in one or two coordinates only, deep trees) and two far clusters
    (one huge leaf-to-leaf gap)."""
    rng = np.random.default_rng(SHAPE_SEEDS[shape])
    n = 4096
    if shape == "duplicates":
        x = np.repeat(rng.random((64, 3)).astype(np.float32) * 3.0, 64, axis=0)
        x = x[rng.permutation(n)]
    elif shape == "line":
        x = np.zeros((n, 3), np.float32)
        x[:, 0] = rng.random(n) * 40.0
    elif shape == "plane":
        x = np.zeros((n, 3), np.float32)
        x[:, :2] = rng.random((n, 2)) * 20.0
    else:
        x = (rng.random((n, 3)) * 2.0).astype(np.float32)
        x[n // 2:] += 500.0
    lo = sc = None
    if morton == 1:
        # the device derives the interleaved key's box from the data's bounding
        # box (scale 1024 / extent, 0 on a flat axis); give the oracle the same
        lo = x.min(axis=0)
        ext = (x.max(axis=0) - lo).astype(np.float32)
        sc = np.where(ext > 0, np.float32(1024.0) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    for pred in (pkg.PRED_SCALAR, pkg.PRED_SIMD):
        _compare_with_oracle(pkg, O, x, 4, 0.3, pred, morton, lo=lo, scale=sc)


@pytest.mark.parametrize("n", [10, 100, 1000, 2000, 3000, 20000])
def test_size_sweep_vs_oracle_and_bruteforce(pkg, O, n):
    x = _rand(n, 1.0, 12 + n)
    got = _compare_with_oracle(pkg, O, x, 5, 0.1, pkg.PRED_SIMD)
    b = pkg.threshold_unique_pairs(x, 0.1)
    assert np.array_equal(rows_canonical(pairs_to_rows(got)), rows_canonical(pairs_to_rows(b)))
    assert np.array_equal(pairs_to_rows(b), pairs_to_rows(O.bruteforce(x, 0.1)))


def test_density08_configs_vs_oracle(pkg, O):
    L = (4096 / 0.8) ** (1 / 3)
    x = _rand(4096, L, 13)
    _compare_with_oracle(pkg, O, x, 4, 2.5, pkg.PRED_SCALAR)
    _compare_with_oracle(pkg, O, x, 4, 2.8, pkg.PRED_SIMD)
    lo = x.min(0)
    sc = (np.float32(1024.0) / (x.max(0) - lo)).astype(np.float32)
    _compare_with_oracle(pkg, O, x, 4, 2.8, pkg.PRED_SCALAR, morton=pkg.MORTON_INTERLEAVE,
                         lo=lo, scale=sc)


def test_margin_pairs_follow_each_predicate(pkg, O):
    r, pts = margin_set()
    spec = pkg.SpheresBVHSpecs(neighbor_distance=r, atom_count=len(pts), atomsperleaf=2)
    for pred in (pkg.PRED_SCALAR, pkg.PRED_SIMD):
        out, n = pkg.neighborlist_device(pts, spec, pred)
        ref, _, _ = O.leafbuild_traverse_bvh(pts, 2, r, predicate=pred)
        assert np.array_equal(pairs_to_rows(out.to_host(n)), pairs_to_rows(ref))
    spec4 = pkg.SpheresBVHSpecs(neighbor_distance=r, atom_count=len(pts), atomsperleaf=4)
    assert len(pkg.leafbuild_traverse_bvh(pts, spec4)) != len(pkg.simdbuild_traverse_bvh(pts, spec4))


def test_tree_deterministic_over_rebuilds(pkg):
    L = (16384 / 0.8) ** (1 / 3)
    x = _rand(16384, L, 14)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.8, atom_count=len(x), atomsperleaf=4)
    first = None
    for _ in range(100):
        pkg.neighborlist_device(x, spec, pkg.PRED_SCALAR, pkg.MORTON_INTERLEAVE)
        nodes, _, perm = pkg.bvh.tree_export(spec, x)
        b = nodes.tobytes() + perm.tobytes()
        if first is None:
            first = b
        assert b == first


def test_capacity_protocol(pkg, O):
    import ctypes
    x = _rand(2000, 1.0, 15)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=0.1, atom_count=len(x), atomsperleaf=4)
    ctx = pkg.default_context()
    d = pkg.DeviceArray.from_host(ctx, x)
    out = pkg.DeviceArray(ctx, (10,), pkg.PAIR_DTYPE)
    cnt = ctypes.c_int64(0)
    rc = pkg.library().nd_leafbuild_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), 0, 0,
                                                  out.ptr, 10, ctypes.byref(cnt))
    ref, _, _ = O.leafbuild_traverse_bvh(x, 4, 0.1)
    assert rc == pkg._lib.ND_ERR_CAPACITY and cnt.value == len(ref)
    out = pkg.DeviceArray(ctx, (cnt.value,), pkg.PAIR_DTYPE)
    rc = pkg.library().nd_leafbuild_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), 0, 0,
                                                  out.ptr, cnt.value, ctypes.byref(cnt))
    assert rc == 0 and np.array_equal(pairs_to_rows(out.to_host(cnt.value)), pairs_to_rows(ref))


def test_host_buffer_entry_point(pkg, O):
    import ctypes
    x = _rand(3000, 1.0, 16)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=0.1, atom_count=len(x), atomsperleaf=5)
    ref, _, _ = O.leafbuild_traverse_bvh(x, 5, 0.1)
    out = np.zeros(len(ref) + 5, pkg.PAIR_DTYPE)
    cnt = ctypes.c_int64(0)
    ctx = pkg.default_context()
    rc = pkg.library().nd_leafbuild_traverse_bvh_host(ctx.handle, x.ctypes.data_as(ctypes.c_void_p),
                                                       ctypes.byref(spec.c), 0, 0,
                                                       out.ctypes.data_as(ctypes.c_void_p),
                                                       len(out), ctypes.byref(cnt))
    assert rc == 0 and cnt.value == len(ref)
    assert np.array_equal(pairs_to_rows(out[: cnt.value]), pairs_to_rows(ref))


def test_full_size_interleave_exact_vs_oracle(pkg, O):
    """N = 262,144 at density 0.8, r = 2.8 (the simulation list radius)."""
    n = 262144
    L = (n / 0.8) ** (1 / 3)
    x = _rand(n, L, 17)
    lo = x.min(0)
    sc = (np.float32(1024.0) / (x.max(0) - lo)).astype(np.float32)
    _compare_with_oracle(pkg, O, x, 4, 2.8, pkg.PRED_SCALAR, morton=pkg.MORTON_INTERLEAVE, lo=lo,
                         scale=sc)


def test_full_size_pair_set_independent_of_key(pkg):
    """Property at N = 262,144: the masked reference key and the interleaved
    key give the same canonical pair set."""
    n = 262144
    L = (n / 0.8) ** (1 / 3)
    x = _rand(n, L, 18)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=n, atomsperleaf=4)
    a, na = pkg.neighborlist_device(x, spec, pkg.PRED_SCALAR, pkg.MORTON_INTERLEAVE)
    ra = rows_canonical(pairs_to_rows(a.to_host(na)))
    b, nb = pkg.neighborlist_device(x, spec, pkg.PRED_SCALAR, pkg.MORTON_MASKED)
    rb = rows_canonical(pairs_to_rows(b.to_host(nb)))
    assert na == nb and np.array_equal(ra, rb)


@pytest.mark.parametrize("apl,pred", [(4, 0), (4, 1), (5, 0), (32, 0)])
def test_async_entry_point_matches_sync(pkg, apl, pred):
    """The stream-ordered one-call list: same pairs in the same order as the
    synchronous call, count and status in device memory, nothing written when
    the buffer is too small, and a domain error reported as status bit 1."""
    import ctypes
    n = 4000
    x = _rand(n, (n / 0.8) ** (1 / 3), 60 + apl)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=n, atomsperleaf=apl)
    ref_out, nref = pkg.neighborlist_device(x, spec, pred, pkg.MORTON_INTERLEAVE)
    ref = pairs_to_rows(ref_out.to_host(nref))
    ctx = pkg.default_context()
    lib = pkg.library()
    d = pkg.DeviceArray.from_host(ctx, x)
    cnt = pkg.DeviceArray.from_host(ctx, np.zeros(1, np.int64))
    st = pkg.DeviceArray.from_host(ctx, np.full(1, 99, np.int32))

    def call(out, cap, xs=d, mode=pkg.MORTON_INTERLEAVE):
        rc = lib.nd_leafbuild_traverse_bvh_async(ctx.handle, xs.ptr, ctypes.byref(spec.c), pred, mode,
                                                 out.ptr, cap, cnt.ptr, st.ptr)
        assert rc == 0
        ctx.synchronize()
        return int(cnt.to_host()[0]), int(st.to_host()[0])

    out = pkg.DeviceArray(ctx, (nref + 16,), pkg.PAIR_DTYPE)
    assert call(out, nref + 16) == (nref, 0)
    assert np.array_equal(pairs_to_rows(out.to_host(nref)), ref)
    sentinel = np.zeros(10, pkg.PAIR_DTYPE)
    sentinel["i"] = -7
    small = pkg.DeviceArray.from_host(ctx, sentinel)
    assert call(small, 10) == (nref, 0)
    assert np.array_equal(small.to_host()["i"], sentinel["i"])
    bad = x.copy()
    bad[3, 1] = 3.0e9  # beyond round(Int32, x / binwidth) in the reference key
    _, status = call(out, nref + 16, pkg.DeviceArray.from_host(ctx, bad), pkg.MORTON_MASKED)
    assert status & 1


def _hip():
    """The HIP runtime the library is linked against (already loaded)."""
    import ctypes
    h = ctypes.CDLL("libamdhip64.so.7")
    P = ctypes.c_void_p
    h.hipStreamBeginCapture.argtypes = [P, ctypes.c_int]
    h.hipStreamEndCapture.argtypes = [P, ctypes.POINTER(P)]
    h.hipGraphInstantiate.argtypes = [ctypes.POINTER(P), P, P, P, ctypes.c_size_t]
    h.hipGraphLaunch.argtypes = [P, P]
    h.hipGraphExecDestroy.argtypes = [P]
    h.hipGraphDestroy.argtypes = [P]
    return h


def test_async_entry_point_in_a_captured_graph(pkg):
    """The stream-ordered one-call list captured once in a hipGraph and
    replayed on new positions written into the same input buffer: every
    replay equals the synchronous call on those positions, also after other
    entry points on the context grew their own work buffers (the async form's
    buffers are its own), with host work and synchronisation between replays."""
    import ctypes
    hip = _hip()
    n, apl = 8000, 4
    L = (n / 0.8) ** (1 / 3)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=n, atomsperleaf=apl)
    ctx = pkg.Context(0)
    lib = pkg.library()
    xs = [_rand(n, L, 900 + k) for k in range(4)]
    d = pkg.DeviceArray.from_host(ctx, xs[0])
    cnt = pkg.DeviceArray.from_host(ctx, np.zeros(1, np.int64))
    st = pkg.DeviceArray.from_host(ctx, np.zeros(1, np.int32))
    cap = 400000
    out = pkg.DeviceArray(ctx, (cap,), pkg.PAIR_DTYPE)
    args = (ctx.handle, d.ptr, ctypes.byref(spec.c), pkg.PRED_SCALAR, pkg.MORTON_INTERLEAVE, out.ptr, cap,
            cnt.ptr, st.ptr)
    assert lib.nd_leafbuild_traverse_bvh_async(*args) == 0      # sizes its work buffers
    ctx.synchronize()
    stream = ctypes.c_void_p()
    assert lib.nd_ctx_stream(ctx.handle, ctypes.byref(stream)) == 0
    graph, gexec = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipStreamBeginCapture(stream, 2) == 0            # relaxed
    assert lib.nd_leafbuild_traverse_bvh_async(*args) == 0
    assert hip.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
    assert hip.hipGraphInstantiate(ctypes.byref(gexec), graph, None, None, 0) == 0
    try:
        for k, x in enumerate(xs):
            if k == 2:
                # other entry points grow the context's shared buffers
                big = _rand(3 * n, (3 * n / 0.8) ** (1 / 3), 77)
                pkg.cells.neighborlist(big, 2.5, ctx=ctx)
                pkg.threshold_unique_pairs(big[:5000], 2.5, ctx=ctx)
            ref_out, nref = pkg.neighborlist_device(x, spec, pkg.PRED_SCALAR, pkg.MORTON_INTERLEAVE, ctx=ctx)
            ref = pairs_to_rows(ref_out.to_host(nref))
            check = ctypes.c_void_p(d.ptr.value)
            assert lib.nd_memcpy_h2d(ctx.handle, check, x.ctypes.data_as(ctypes.c_void_p), x.nbytes) == 0
            assert hip.hipGraphLaunch(gexec, stream) == 0
            ctx.synchronize()
            assert int(cnt.to_host()[0]) == nref and int(st.to_host()[0]) == 0, k
            assert np.array_equal(pairs_to_rows(out.to_host(nref)), ref), k
    finally:
        hip.hipGraphExecDestroy(gexec)
        hip.hipGraphDestroy(graph)
        ctx.close()
