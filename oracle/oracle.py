"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py.  It restates the reference's neighbour-search
path (refcpu.c) and an FP64 textbook physics oracle (oracle64.c).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,)),
                       ("left", "<i4"), ("skip", "<i4")])
PAIR_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("d", "<f4")])

PRED_SCALAR, PRED_SIMD = 0, 1
MORTON_MASKED, MORTON_INTERLEAVE, MORTON_HILBERT = 0, 1, 2
QUERY_OWN_LEAF, QUERY_ROOT = 0, 1


class Adj64Params(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("sigma", ctypes.c_double), ("rc", ctypes.c_double),
                ("k_coul", ctypes.c_double), ("rc_coul", ctypes.c_double), ("dt", ctypes.c_double),
                ("rdf_rmax", ctypes.c_double), ("rdf_h", ctypes.c_double), ("volume", ctypes.c_double),
                ("nbins", ctypes.c_int), ("nsteps", ctypes.c_int)]


class O64Params(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("sigma", ctypes.c_double),
                ("rc_lj", ctypes.c_double), ("k_coul", ctypes.c_double),
                ("rc_coul", ctypes.c_double), ("shift_lj", ctypes.c_int32),
                ("shift_coul", ctypes.c_int32)]


class BaseParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_float), ("sigma", ctypes.c_float), ("rc", ctypes.c_float),
                ("k_coul", ctypes.c_float), ("rc_coul", ctypes.c_float),
                ("dt", ctypes.c_float), ("skin", ctypes.c_float), ("lo", ctypes.c_float),
                ("hi", ctypes.c_float), ("apl", ctypes.c_int32),
                ("rebuild_every", ctypes.c_int32), ("walls", ctypes.c_int32),
                ("predicate", ctypes.c_int32), ("morton_mode", ctypes.c_int32),
                ("nthreads", ctypes.c_int32)]


class TreeOut(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("perm", ctypes.c_void_p),
                ("codes_sorted", ctypes.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.ref_binwidth.restype = f32
        L.ref_morton_masked.argtypes = [P, i64, P]
        L.ref_morton_interleave.argtypes = [P, i64, P, P, P]
        L.ref_morton_hilbert.argtypes = [P, i64, P, P, P]
        L.ref_hilbert_cells.argtypes = [P, i64, P]
        L.ref_hilbert_cells.restype = None
        L.ref_sortperm.argtypes = [P, i64, P]
        L.ref_sortperm.restype = None
        L.ref_leaf_boxes.argtypes = [P, i32, i32, f32, P]
        L.ref_leaf_boxes.restype = None
        L.ref_delta_public.argtypes = [i32, P, i32, i32]
        L.ref_delta_public.restype = i32
        L.ref_build_tree.argtypes = [P, P, i32, i32, f32, P, ctypes.c_int, ctypes.c_int]
        L.ref_traverse.argtypes = [P, P, P, i32, i32, f32, ctypes.c_int, P, i64,
                                   ctypes.c_int, P]
        L.ref_traverse.restype = i64
        L.ref_leafbuild_traverse_bvh.argtypes = [P, i32, i32, f32, ctypes.c_int, ctypes.c_int,
                                                 P, P, P, i64, ctypes.c_int, P, P]
        L.ref_leafbuild_traverse_bvh.restype = i64
        L.ref_bruteforce.argtypes = [P, i32, f32, P, i64]
        L.ref_bruteforce.restype = i64
        L.ref_cluster_boxes.argtypes = [P, i32, i32, f32, P]
        L.ref_cluster_boxes.restype = None
        L.ref_neighbor_traverse.argtypes = [P, P, P, i32, i32, f32, ctypes.c_int, P, i64,
                                            ctypes.c_int]
        L.ref_neighbor_traverse.restype = i64
        L.ref_build_traverse_bvh.argtypes = [P, i32, i32, f32, ctypes.c_int, ctypes.c_int, P, P,
                                             P, i64, ctypes.c_int, P, P]
        L.ref_build_traverse_bvh.restype = i64
        L.ref_quantized_positions.argtypes = [P, i32, P]
        L.ref_quantized_positions.restype = None
        L.o64_forces.argtypes = [P, P, i32, P, P, P, P]
        L.o64_forces.restype = i64
        L.o64_forces_pairs.argtypes = [P, P, i32, P, P, i64, P, P, P, P]
        L.o64_forces_pairs.restype = None
        L.o64_vv_run.argtypes = [P, P, P, P, i32, P, ctypes.c_double, i32, i32,
                                 ctypes.c_double, ctypes.c_double, P, P]
        L.o64_vv_run_thermo.argtypes = [P, P, P, P, i32, P, ctypes.c_double, i32, i32,
                                        ctypes.c_double, ctypes.c_double, i32, ctypes.c_double,
                                        ctypes.c_double, P, P]
        L.base_run.argtypes = [P, P, P, i32, P, i32, P, P]
        L.adj64_run.argtypes = [P, P, P, ctypes.c_int, P, P, P, P, P, P, P, P]
        L.adj64_forces.argtypes = [P, P, ctypes.c_int, P, P, P]
        L.adj64_forces.restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def binwidth():
    return np.float32(lib().ref_binwidth())


def morton_masked(xyz):
    xyz = _f32(xyz).reshape(-1, 3)
    out = np.empty(len(xyz), np.int32)
    rc = lib().ref_morton_masked(_p(xyz), len(xyz), _p(out))
    if rc != 0:
        raise ValueError("oracle: coordinate outside the Int32 quantisation domain")
    return out


def interleave_scale(lo, hi):
    lo = np.asarray(lo, np.float32).reshape(3)
    hi = np.asarray(hi, np.float32).reshape(3)
    return lo, (np.float32(1024.0) / (hi - lo)).astype(np.float32)


def morton_interleave(xyz, lo, scale):
    xyz = _f32(xyz).reshape(-1, 3)
    lo, scale = _f32(lo), _f32(scale)
    out = np.empty(len(xyz), np.int32)
    rc = lib().ref_morton_interleave(_p(xyz), len(xyz), _p(lo), _p(scale), _p(out))
    if rc != 0:
        raise ValueError("oracle: non-finite coordinate")
    return out


def hilbert_cells(cells):
    """Hilbert index of integer cells (n x 3, each in [0, 1024))."""
    c = np.ascontiguousarray(cells, np.uint32).reshape(-1, 3)
    out = np.empty(len(c), np.int32)
    lib().ref_hilbert_cells(_p(c), len(c), _p(out))
    return out


def morton_hilbert(xyz, lo, scale):
    """Hilbert key of the interleaved key's 10-bit cells."""
    xyz = _f32(xyz).reshape(-1, 3)
    lo, scale = _f32(lo), _f32(scale)
    out = np.empty(len(xyz), np.int32)
    rc = lib().ref_morton_hilbert(_p(xyz), len(xyz), _p(lo), _p(scale), _p(out))
    if rc != 0:
        raise ValueError("oracle: non-finite coordinate")
    return out


def sortperm(codes):
    codes = np.ascontiguousarray(codes, np.int32)
    perm = np.empty(len(codes), np.int32)
    lib().ref_sortperm(_p(codes), len(codes), _p(perm))
    return perm


def delta(i, codes_sorted, nL, apl):
    codes_sorted = np.ascontiguousarray(codes_sorted, np.int32)
    return int(lib().ref_delta_public(int(i), _p(codes_sorted), int(nL), int(apl)))


def build_tree(xyz_sorted, codes_sorted, apl, r, override_root=True, nthreads=1):
    xyz_sorted = _f32(xyz_sorted).reshape(-1, 3)
    codes_sorted = np.ascontiguousarray(codes_sorted, np.int32)
    N = len(xyz_sorted)
    nL = N // apl
    nodes = np.zeros(2 * nL - 1, NODE_DTYPE)
    rc = lib().ref_build_tree(_p(xyz_sorted), _p(codes_sorted), N, apl, np.float32(r),
                              _p(nodes), int(override_root), int(nthreads))
    if rc != 0:
        raise ValueError("oracle: bad BVH specification")
    return nodes


def traverse(nodes, xyz_sorted, idx_sorted, apl, r, predicate=PRED_SCALAR, nthreads=1):
    xyz_sorted = _f32(xyz_sorted).reshape(-1, 3)
    idx_sorted = np.ascontiguousarray(idx_sorted, np.int32)
    N = len(xyz_sorted)
    tests = ctypes.c_int64(0)
    n = lib().ref_traverse(_p(nodes), _p(xyz_sorted), _p(idx_sorted), N, apl, np.float32(r),
                           predicate, None, 0, nthreads, ctypes.byref(tests))
    out = np.empty(n, PAIR_DTYPE)
    n2 = lib().ref_traverse(_p(nodes), _p(xyz_sorted), _p(idx_sorted), N, apl, np.float32(r),
                            predicate, _p(out), n, nthreads, None)
    assert n2 == n
    return out, tests.value


def leafbuild_traverse_bvh(xyz, apl, r, predicate=PRED_SCALAR, morton=MORTON_MASKED,
                           lo=None, scale=None, nthreads=1, want_tree=False):
    """One-call neighbour list: returns (pairs, tree dict or None, timing[3])."""
    xyz = _f32(xyz).reshape(-1, 3)
    N = len(xyz)
    if N % apl != 0 or N // apl < 2:
        raise ValueError("bad BVH specification")
    nL = N // apl
    lo = _f32(lo if lo is not None else np.zeros(3))
    scale = _f32(scale if scale is not None else np.ones(3))
    timing = np.zeros(3, np.float64)
    n = lib().ref_leafbuild_traverse_bvh(_p(xyz), N, apl, np.float32(r), predicate, morton,
                                         _p(lo), _p(scale), None, 0, nthreads, None,
                                         _p(timing))
    if n < 0:
        raise ValueError("oracle: error %d" % n)
    out = np.empty(n, PAIR_DTYPE)
    tree = None
    tout = None
    if want_tree:
        tree = {"nodes": np.zeros(2 * nL - 1, NODE_DTYPE), "perm": np.zeros(N, np.int32),
                "codes_sorted": np.zeros(N, np.int32)}
        tout = TreeOut(_p(tree["nodes"]).value, _p(tree["perm"]).value,
                       _p(tree["codes_sorted"]).value)
    n2 = lib().ref_leafbuild_traverse_bvh(_p(xyz), N, apl, np.float32(r), predicate, morton,
                                          _p(lo), _p(scale), _p(out), n, nthreads,
                                          ctypes.byref(tout) if tout is not None else None,
                                          _p(timing))
    assert n2 == n
    return out, tree, timing


def cluster_boxes(xyz_sorted, apl, r):
    """cluster_primitives: leaf boxes of +- r (nL nodes, left = skip = 0)."""
    xyz_sorted = _f32(xyz_sorted).reshape(-1, 3)
    nL = len(xyz_sorted) // apl
    nodes = np.zeros(nL, NODE_DTYPE)
    lib().ref_cluster_boxes(_p(xyz_sorted), nL, apl, np.float32(r), _p(nodes))
    return nodes


def neighbor_traverse(nodes, xyz_sorted, idx_sorted, apl, r, mode=QUERY_OWN_LEAF, nthreads=1):
    """Point-query walk over a given tree (neighbor_traverse / the root form)."""
    xyz_sorted = _f32(xyz_sorted).reshape(-1, 3)
    idx_sorted = np.ascontiguousarray(idx_sorted, np.int32)
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    N = len(xyz_sorted)
    args = (_p(nodes), _p(xyz_sorted), _p(idx_sorted), N, apl, np.float32(r), mode)
    n = lib().ref_neighbor_traverse(*args, None, 0, nthreads)
    if n < 0:
        raise ValueError("oracle: bad BVH specification")
    out = np.empty(n, PAIR_DTYPE)
    assert lib().ref_neighbor_traverse(*args, _p(out), n, nthreads) == n
    return out


def build_traverse_bvh(xyz, apl, r, mode=QUERY_OWN_LEAF, morton=MORTON_MASKED, lo=None,
                       scale=None, nthreads=1):
    """TreeData + neighbor_traverse: returns (pairs, tree dict with store)."""
    xyz = _f32(xyz).reshape(-1, 3)
    N = len(xyz)
    if N % apl != 0 or N // apl < 2:
        raise ValueError("bad BVH specification")
    nL = N // apl
    lo = _f32(lo if lo is not None else np.zeros(3))
    scale = _f32(scale if scale is not None else np.ones(3))
    tree = {"nodes": np.zeros(2 * nL - 1, NODE_DTYPE), "perm": np.zeros(N, np.int32),
            "codes_sorted": np.zeros(N, np.int32), "store": np.zeros(nL - 1, np.int32)}
    tout = TreeOut(_p(tree["nodes"]).value, _p(tree["perm"]).value, _p(tree["codes_sorted"]).value)
    head = (_p(xyz), N, apl, np.float32(r), mode, morton, _p(lo), _p(scale))
    n = lib().ref_build_traverse_bvh(*head, None, 0, nthreads, None, None)
    if n < 0:
        raise ValueError("oracle: error %d" % n)
    out = np.empty(n, PAIR_DTYPE)
    n2 = lib().ref_build_traverse_bvh(*head, _p(out), n, nthreads, ctypes.byref(tout),
                                      _p(tree["store"]))
    assert n2 == n
    return out, tree


def quantized_positions(xyz):
    """quantized_positions!: (N, 3) int32, 1-based rank of each atom per axis."""
    xyz = _f32(xyz).reshape(-1, 3)
    q = np.zeros((len(xyz), 3), np.int32)
    lib().ref_quantized_positions(_p(xyz), len(xyz), _p(q))
    return q


def bruteforce(xyz, r):
    xyz = _f32(xyz).reshape(-1, 3)
    n = lib().ref_bruteforce(_p(xyz), len(xyz), np.float32(r), None, 0)
    out = np.empty(n, PAIR_DTYPE)
    lib().ref_bruteforce(_p(xyz), len(xyz), np.float32(r), _p(out), n)
    return out


def canonical(pairs):
    """(min(i,j), max(i,j), d-bits) rows sorted lexicographically."""
    i = np.minimum(pairs["i"], pairs["j"]).astype(np.int64)
    j = np.maximum(pairs["i"], pairs["j"]).astype(np.int64)
    d = pairs["d"].view(np.uint32).astype(np.int64)
    a = np.stack([i, j, d], axis=1)
    order = np.lexsort((a[:, 2], a[:, 1], a[:, 0]))
    return a[order]


def o64_params(eps=1.0, sigma=1.0, rc_lj=2.5, k_coul=1.0, rc_coul=0.0, shift_lj=False,
               shift_coul=False):
    return O64Params(eps, sigma, rc_lj, k_coul, rc_coul, int(shift_lj), int(shift_coul))


def forces64(x, q, params):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    N = len(x)
    qq = None if q is None else np.ascontiguousarray(q, np.float64)
    F = np.zeros((N, 3), np.float64)
    U = ctypes.c_double(0)
    W = ctypes.c_double(0)
    npairs = lib().o64_forces(_p(x), _p(qq), N, ctypes.byref(params), _p(F), ctypes.byref(U),
                              ctypes.byref(W))
    return F, U.value, W.value, npairs


def forces64_pairs(x, q, pi, pj, params):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    N = len(x)
    qq = None if q is None else np.ascontiguousarray(q, np.float64)
    pi = np.ascontiguousarray(pi, np.int32)
    pj = np.ascontiguousarray(pj, np.int32)
    F = np.zeros((N, 3), np.float64)
    U = ctypes.c_double(0)
    W = ctypes.c_double(0)
    lib().o64_forces_pairs(_p(x), _p(qq), N, _p(pi), _p(pj), len(pi), ctypes.byref(params),
                           _p(F), ctypes.byref(U), ctypes.byref(W))
    return F, U.value, W.value


def vv_run64(x, v, q, m, params, dt, nsteps, walls=False, lo=0.0, hi=0.0, thermostat_every=0,
             temperature=1.0, gamma=1.0):
    """FP64 kick-drift-force-kick velocity Verlet (o64_vv_run_thermo): returns
    the final (x, v, F) and the (nsteps + 1) x 3 rows (kinetic, potential,
    virial) of every step; thermostat_every > 0 rescales the velocities as
    rescale_velocity! does (Simulator.jl:119-144, T from v^2)."""
    x = np.array(x, np.float64).reshape(-1, 3)
    v = np.array(v, np.float64).reshape(-1, 3)
    N = len(x)
    qq = None if q is None else np.ascontiguousarray(q, np.float64)
    m = np.ascontiguousarray(m if m is not None else np.ones(N), np.float64)
    F = np.zeros((N, 3), np.float64)
    E = np.zeros((nsteps + 1, 3), np.float64)
    lib().o64_vv_run_thermo(_p(x), _p(v), _p(qq), _p(m), N, ctypes.byref(params), dt, nsteps,
                            int(walls), lo, hi, int(thermostat_every), temperature, gamma, _p(F),
                            _p(E))
    return x, v, F, E


def base_run(x, v, q, params, nsteps):
    """Threaded FP32 CPU baseline; returns (x, v, times[4], pairs_evaluated)."""
    x = np.array(x, np.float32).reshape(-1, 3)
    v = np.array(v, np.float32).reshape(-1, 3)
    qq = None if q is None else np.ascontiguousarray(q, np.float32)
    times = np.zeros(4, np.float64)
    pe = ctypes.c_int64(0)
    rc = lib().base_run(_p(x), _p(v), _p(qq), len(x), ctypes.byref(params), nsteps, _p(times),
                        ctypes.byref(pe))
    if rc != 0:
        raise ValueError("oracle baseline error %d" % rc)
    return x, v, times, pe.value


def adj64_run(x0, v0, q, params, target=None, grads=False):
    """FP64 CPU oracle of the differentiable simulation (oracle/adjoint64.c):
    returns (loss, rdf) or, with grads=True, (loss, rdf, dx0, dv0, dq, d_eps_sigma)."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    v0 = np.ascontiguousarray(v0, np.float64).reshape(-1, 3)
    n = len(x0)
    qq = None if q is None else np.ascontiguousarray(q, np.float64)
    tg = None if target is None else np.ascontiguousarray(target, np.float64)
    g = np.zeros(params.nbins)
    L = ctypes.c_double(0)
    dx = np.zeros((n, 3)) if grads else None
    dv = np.zeros((n, 3)) if grads else None
    dq = np.zeros(n) if grads else None
    dp = np.zeros(2) if grads else None
    rc = lib().adj64_run(_p(x0), _p(v0), _p(qq), n, ctypes.byref(params), _p(tg), _p(g),
                         ctypes.byref(L), _p(dx), _p(dv), _p(dq), _p(dp))
    if rc != 0:
        raise MemoryError("adj64_run")
    if grads:
        return L.value, g, dx, dv, dq, dp
    return L.value, g


def adj64_forces(x, q, params):
    """Forces and shifted-force energy of the adjoint oracle's forward model."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    n = len(x)
    qq = None if q is None else np.ascontiguousarray(q, np.float64)
    F = np.zeros((n, 3))
    U = ctypes.c_double(0)
    lib().adj64_forces(_p(x), _p(qq), n, ctypes.byref(params), _p(F), ctypes.byref(U))
    return F, U.value
