"""GPU: the uniform cell-list search (CellListMap's neighborlist(x, cutoff), the
comparison point of assets/BVHBenchSuite.jl:307) against brute force, bit for
bit: the canonical pair set and every distance's bits, for both predicates,
degenerate inputs, non-finite coordinates, the capacity protocol and, at
N = 262,144, against the GPU brute force; and the simulation's cell-list walk
(list_walk 3) against the LBVH walks and the FP64 oracle."""
import ctypes

import numpy as np
import pytest
from conftest import bruteforce_np, pairs_to_rows, rows_canonical
from test_oracle_neighbor import margin_set

pytestmark = pytest.mark.gpu


def _rand(n, L, seed, lo=0.0):
    return (np.random.default_rng(seed).random((n, 3)) * L + lo).astype(np.float32)


def _cells(pkg, x, r, pred):
    return rows_canonical(pairs_to_rows(pkg.cell_neighborlist(x, r, pred)))


@pytest.mark.parametrize("n", [0, 1, 2, 10, 1000, 5000])
def test_matches_bruteforce(pkg, O, n):
    L = max(n / 0.8, 1.0) ** (1 / 3)
    x = _rand(n, L, 40 + n)
    got = pkg.cell_neighborlist(x, 2.5)
    if n >= 2:
        assert np.array_equal(rows_canonical(pairs_to_rows(got)), pairs_to_rows(O.bruteforce(x, 2.5)))
    else:
        assert len(got) == 0
    # canonical orientation on output: i < j
    assert np.all(got["i"] < got["j"])


@pytest.mark.parametrize("r", [1e-3, 0.05, 0.3, 2.0])
def test_scalar_predicate_vs_definition(pkg, r):
    x = _rand(3000, 1.0, 41)
    assert np.array_equal(_cells(pkg, x, r, pkg.PRED_SCALAR), bruteforce_np(x, r, 0))
    assert np.array_equal(_cells(pkg, x, r, pkg.PRED_SIMD), bruteforce_np(x, r, 1))


def test_margin_pairs_follow_each_predicate(pkg):
    r, pts = margin_set()
    a = _cells(pkg, pts, r, pkg.PRED_SCALAR)
    b = _cells(pkg, pts, r, pkg.PRED_SIMD)
    assert np.array_equal(a, bruteforce_np(pts, r, 0))
    assert np.array_equal(b, bruteforce_np(pts, r, 1))
    assert len(a) != len(b)


# fixed seeds, one per input shape (reproducible in every process)
SHAPE_SEEDS = {"duplicates": 201, "line": 202, "plane": 203, "two_clusters": 204, "one_point": 205,
               "wide_sparse": 206}


@pytest.mark.parametrize("shape", ["duplicates", "line", "plane", "two_clusters", "one_point",
                                   "wide_sparse"])
def test_degenerate_inputs(pkg, O, shape):
    """Collisions (pairs at distance 0, one crowded cell), flat axes (one cell
    across), a huge gap (the cell count cap enlarges cells), all atoms on one
    point (zero extent) and a box 10^4 cutoffs wide (1024 cells per axis cap)."""
    rng = np.random.default_rng(SHAPE_SEEDS[shape])
    n = 2048
    if shape == "duplicates":
        x = np.repeat(rng.random((32, 3)).astype(np.float32) * 3.0, 64, axis=0)[rng.permutation(n)]
    elif shape == "line":
        x = np.zeros((n, 3), np.float32)
        x[:, 0] = rng.random(n) * 40.0
    elif shape == "plane":
        x = np.zeros((n, 3), np.float32)
        x[:, :2] = rng.random((n, 2)) * 20.0
    elif shape == "two_clusters":
        x = (rng.random((n, 3)) * 2.0).astype(np.float32)
        x[n // 2:] += 500.0
    elif shape == "one_point":
        x = np.full((300, 3), 1.25, np.float32)
    else:
        x = (rng.random((n, 3)) * 3000.0).astype(np.float32)
    r = 0.3 if shape != "wide_sparse" else 60.0
    ref = pairs_to_rows(O.bruteforce(x, r))
    assert np.array_equal(_cells(pkg, x, r, pkg.PRED_SIMD), ref)
    assert np.array_equal(_cells(pkg, x, r, pkg.PRED_SCALAR), bruteforce_np(x, r, 0))


def test_nonfinite_atoms_pair_with_nothing(pkg, O):
    x = _rand(2000, 10.0, 42)
    x[5] = [np.nan, 1.0, 1.0]
    x[77, 2] = np.inf
    x[1500] = -np.inf
    got = _cells(pkg, x, 1.0, pkg.PRED_SIMD)
    assert np.array_equal(got, pairs_to_rows(O.bruteforce(x, 1.0)))
    assert not np.isin(got[:, :2], [5, 77, 1500]).any()


def test_bad_arguments_are_errors(pkg):
    x = _rand(100, 1.0, 43)
    for r in (0.0, -1.0, np.inf, np.nan):
        with pytest.raises(pkg.NDError):
            pkg.cell_neighborlist(x, r)
    with pytest.raises(pkg.SpecError):
        pkg.cell_neighborlist(x, 0.1, predicate=7)


def test_capacity_protocol_and_deterministic_order(pkg, O):
    x = _rand(2000, 1.0, 44)
    ctx = pkg.default_context()
    d = pkg.DeviceArray.from_host(ctx, x)
    out = pkg.DeviceArray(ctx, (10,), pkg.PAIR_DTYPE)
    cnt = ctypes.c_int64(0)
    lib = pkg.library()
    rc = lib.nd_cell_neighborlist(ctx.handle, d.ptr, len(x), np.float32(0.1), pkg.PRED_SIMD, out.ptr, 10,
                                  ctypes.byref(cnt))
    ref = pairs_to_rows(O.bruteforce(x, 0.1))
    assert rc == pkg._lib.ND_ERR_CAPACITY and cnt.value == len(ref)
    out = pkg.DeviceArray(ctx, (cnt.value,), pkg.PAIR_DTYPE)
    rows = []
    for _ in range(3):
        rc = lib.nd_cell_neighborlist(ctx.handle, d.ptr, len(x), np.float32(0.1), pkg.PRED_SIMD, out.ptr,
                                      cnt.value, ctypes.byref(cnt))
        assert rc == 0
        rows.append(pairs_to_rows(out.to_host(cnt.value)))
    assert np.array_equal(rows_canonical(rows[0]), ref)
    assert all(np.array_equal(rows[0], r) for r in rows[1:])


def test_host_buffer_entry_point(pkg, O):
    x = _rand(3000, 1.0, 48)
    ref = pairs_to_rows(O.bruteforce(x, 0.1))
    out = np.zeros(len(ref) + 5, pkg.PAIR_DTYPE)
    cnt = ctypes.c_int64(0)
    ctx = pkg.default_context()
    rc = pkg.library().nd_cell_neighborlist_host(ctx.handle, x.ctypes.data_as(ctypes.c_void_p), len(x),
                                                 np.float32(0.1), pkg.PRED_SIMD,
                                                 out.ctypes.data_as(ctypes.c_void_p), len(out),
                                                 ctypes.byref(cnt))
    assert rc == 0 and cnt.value == len(ref)
    assert np.array_equal(rows_canonical(pairs_to_rows(out[: cnt.value])), ref)


def test_full_size_equals_gpu_bruteforce(pkg):
    """Property at N = 262,144 (density 0.8, r = 2.8): the cell list and the
    O(N^2) brute-force kernel give the same canonical set and distance bits."""
    n = 262144
    L = (n / 0.8) ** (1 / 3)
    x = _rand(n, L, 45)
    a = _cells(pkg, x, 2.8, pkg.PRED_SIMD)
    b = rows_canonical(pairs_to_rows(pkg.threshold_unique_pairs(x, 2.8)))
    assert len(a) > 9_000_000 and np.array_equal(a, b)


# ---- the simulation's cell-list walk ----

def _cfg(pkg, n, seed):
    L = pkg.synth.box_side(n)
    return (pkg.synth.lattice_positions(n, L, seed), pkg.synth.gaussian_velocities(n, 1.0, seed),
            pkg.synth.balanced_charges(n, seed), L)


def _spec(pkg, L, precision="fp32", steps=10, **kw):
    d = dict(duration=steps, stepwidth=0.001, rc=2.5, rc_coul=2.5, skin=0.3, rebuild_every=10, lo=0.0,
             hi=L, precision=precision, energy_every=1)
    d.update(kw)
    return pkg.SimSpec(**d)


def test_cell_walk_gives_the_lbvh_list(pkg):
    x, v, q, L = _cfg(pkg, 32768, 46)
    res = []
    for walk in (0, 3):
        sim = pkg.Simulation(len(x), _spec(pkg, L, list_walk=walk, atomsperleaf=16))
        sim.set_state(x, v, q)
        sim.run(0)
        st = sim.stats()
        res.append((st["list_entries"], st["max_neighbors"], sim.get_state()[2]))
        sim.close()
    assert res[0][:2] == res[1][:2] and res[0][0] > 0
    f0, f3 = res[0][2], res[1][2]
    assert np.linalg.norm(f3 - f0) / np.linalg.norm(f0) < 1e-6


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("mixed", 1e-7), ("fp64", 1e-10)])
@pytest.mark.parametrize("resort_every", [0, 50])
def test_cell_walk_trajectory_vs_oracle(pkg, O, precision, xtol, resort_every):
    x, v, q, L = _cfg(pkg, 4096, 47)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, steps=100, list_walk=3, resort_every=resort_every))
    sim.set_state(x, v, q)
    sim.run(100)
    xg, _, fg = sim.get_state()
    xr, _, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 100,
                              walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < xtol
    assert np.linalg.norm(fg - fr) / np.linalg.norm(fr) < (5e-4 if precision == "fp32" else 1e-5)
    assert sim.stats()["rebuilds"] == 11
    sim.close()


def test_cell_walk_needs_a_box(pkg):
    with pytest.raises(pkg.SpecError):
        pkg.Simulation(64, pkg.SimSpec(duration=0, lo=0.0, hi=0.0, walls=False, morton=pkg.MORTON_MASKED,
                                       list_walk=3))
