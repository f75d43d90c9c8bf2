"""LBVH neighbour search with the reference's entry points.

Mirrors src/Neighbors/BVHTraverse.jl, ext/NaiveSIMD.jl, ext/NaiveKA.jl and
src/Neighbors/AllToAll.jl of NaiveDynamics.jl; every call runs on the GPU
through libnaivehip.so.  Julia's mutating `name!` functions keep their names
without the `!`.  Indices are 0-based (the Julia reference is 1-based).
Pair lists are numpy structured arrays with fields (i, j, d), the layout of
Vector{Tuple{Int32,Int32,Float32}}.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import (MORTON_INTERLEAVE, MORTON_MASKED, NODE_DTYPE, PAIR_DTYPE, PRED_SCALAR,
                   PRED_SIMD, QUERY_OWN_LEAF, QUERY_ROOT, BvhSpec, CapacityError, DeviceArray,
                   SpecError, check, load, ptr)

binwidth = np.float32(1.0 / 1023.0)  # BVHTraverse.jl:184


class SpheresBVHSpecs:
    """SpheresBVHSpecs(; neighbor_distance, atom_count, floattype, atomsperleaf)
    (BVHTraverse.jl:63-94).  Raises SpecError where the reference calls error()."""

    def __init__(self, *, neighbor_distance, atom_count, atomsperleaf, floattype=np.float32):
        if np.dtype(floattype) != np.float32:
            raise SpecError(_lib.ND_ERR_ARG, "only Float32 / Int32 specifications are implemented")
        if int(atom_count) % int(atomsperleaf) != 0:
            raise SpecError(_lib.ND_ERR_ARG, "Please use an 'atomsperleaf' that evenly divides "
                            "into 'atom_count' in BVH Specification")
        if int(atom_count) // int(atomsperleaf) < 2:
            raise SpecError(_lib.ND_ERR_ARG, "Please use more than one leaf in BVH Specification")
        c = BvhSpec()
        rc = load().nd_bvh_spec_init(ctypes.byref(c), np.float32(neighbor_distance),
                                     int(atom_count), int(atomsperleaf))
        check(rc, None, "SpheresBVHSpecs")
        self.c = c

    neighbor_distance = property(lambda s: np.float32(s.c.neighbor_distance))
    atom_count = property(lambda s: s.c.atom_count)
    leaves_count = property(lambda s: s.c.leaves_count)
    branches_count = property(lambda s: s.c.branches_count)
    atomsperleaf = property(lambda s: s.c.atomsperleaf)

    def __repr__(self):
        return ("SpheresBVHSpecs(neighbor_distance=%r, atom_count=%d, leaves_count=%d, "
                "branches_count=%d, atomsperleaf=%d)" % (
                    float(self.neighbor_distance), self.atom_count, self.leaves_count,
                    self.branches_count, self.atomsperleaf))


def _ctx(ctx):
    return ctx if ctx is not None else _lib.default_context()


def _dev_xyz(ctx, position):
    if isinstance(position, DeviceArray):
        return position, False
    x = np.ascontiguousarray(position, dtype=np.float32).reshape(-1, 3)
    return DeviceArray.from_host(ctx, x), True


def mortoncodes(position, spec=None, mode=MORTON_MASKED, lo=None, scale=None, ctx=None):
    """mortoncodes! (BVHTraverse.jl:257-288) -> int32 codes in input order."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    n = d.shape[0]
    out = DeviceArray(ctx, (n,), np.int32)
    lo_p = scale_p = None
    if lo is not None:
        lo = np.ascontiguousarray(lo, np.float32).reshape(3)
        scale = np.ascontiguousarray(scale, np.float32).reshape(3)
        lo_p, scale_p = ptr(lo), ptr(scale)
    rc = load().nd_mortoncodes(ctx.handle, d.ptr, n, mode, lo_p, scale_p, out.ptr)
    check(rc, ctx, "mortoncodes!")
    return out.to_host()


def sortperm(codes, ctx=None):
    """sortperm(codes) (BVHTraverse.jl:570): stable; returns (sorted codes, perm)."""
    ctx = _ctx(ctx)
    c = DeviceArray.from_host(ctx, np.ascontiguousarray(codes, np.int32))
    n = c.shape[0]
    cs = DeviceArray(ctx, (n,), np.int32)
    pm = DeviceArray(ctx, (n,), np.int32)
    check(load().nd_sortperm(ctx.handle, c.ptr, n, cs.ptr, pm.ptr), ctx, "sortperm")
    return cs.to_host(), pm.to_host()


def bounding_volume_hierarchy(xyz_sorted, codes_sorted, spec, override_root=True, ctx=None):
    """leafcluster_primitives + bounding_volume_hierarchy! (BVHTraverse.jl:376-410,
    849-860) on already sorted positions/codes -> nodes (2nL-1, GridKey layout)."""
    ctx = _ctx(ctx)
    x = DeviceArray.from_host(ctx, np.ascontiguousarray(xyz_sorted, np.float32).reshape(-1, 3))
    c = DeviceArray.from_host(ctx, np.ascontiguousarray(codes_sorted, np.int32))
    nodes = DeviceArray(ctx, (2 * spec.leaves_count - 1,), NODE_DTYPE)
    rc = load().nd_bounding_volume_hierarchy(ctx.handle, x.ptr, c.ptr, ctypes.byref(spec.c),
                                             int(override_root), nodes.ptr)
    check(rc, ctx, "bounding_volume_hierarchy!")
    return nodes.to_host()


def leafneighbor_traverse(keys, xyz_sorted, index_sorted, spec, predicate=PRED_SCALAR, ctx=None):
    """leafneighbor_traverse (BVHTraverse.jl:1236-1323) over a given tree."""
    ctx = _ctx(ctx)
    nodes = DeviceArray.from_host(ctx, np.ascontiguousarray(keys, NODE_DTYPE))
    x = DeviceArray.from_host(ctx, np.ascontiguousarray(xyz_sorted, np.float32).reshape(-1, 3))
    ix = DeviceArray.from_host(ctx, np.ascontiguousarray(index_sorted, np.int32))
    count = ctypes.c_int64(0)
    rc = load().nd_leafneighbor_traverse(ctx.handle, nodes.ptr, x.ptr, ix.ptr,
                                         ctypes.byref(spec.c), predicate, None, 0,
                                         ctypes.byref(count))
    if rc == _lib.ND_ERR_CAPACITY or count.value == 0:
        out = DeviceArray(ctx, (max(count.value, 1),), PAIR_DTYPE)
        rc = load().nd_leafneighbor_traverse(ctx.handle, nodes.ptr, x.ptr, ix.ptr,
                                             ctypes.byref(spec.c), predicate, out.ptr,
                                             count.value, ctypes.byref(count))
        check(rc, ctx, "leafneighbor_traverse", count.value)
        return out.to_host(count.value)
    check(rc, ctx, "leafneighbor_traverse", count.value)
    return np.empty(0, PAIR_DTYPE)


class _PairBuffer:
    """Grow-only device output buffer per context (capacity protocol).  Keyed
    by the context object itself (weakly), so a buffer never outlives or
    changes hands between contexts; a buffer freed by its context's close is
    replaced."""

    def __init__(self):
        self.buf = weakref.WeakKeyDictionary()

    def get(self, ctx, cap):
        d = self.buf.get(ctx)
        if d is None or d.ptr is None or d.shape[0] < cap:
            d = DeviceArray(ctx, (cap,), PAIR_DTYPE)
            self.buf[ctx] = d
        return d


_pairbuf = _PairBuffer()


def neighborlist_device(position, spec, predicate=PRED_SCALAR, morton=MORTON_MASKED, ctx=None):
    """One-call neighbour list on device; returns (DeviceArray of pairs, count)."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    if d.shape[0] != spec.atom_count:
        raise SpecError(_lib.ND_ERR_ARG, "position length does not match atom_count")
    cap = max(1024, 48 * spec.atom_count)
    out = _pairbuf.get(ctx, cap)
    count = ctypes.c_int64(0)
    rc = load().nd_leafbuild_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), predicate,
                                          morton, out.ptr, out.shape[0], ctypes.byref(count))
    if rc == _lib.ND_ERR_CAPACITY:
        out = _pairbuf.get(ctx, count.value + count.value // 8 + 1024)
        rc = load().nd_leafbuild_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), predicate,
                                              morton, out.ptr, out.shape[0], ctypes.byref(count))
    check(rc, ctx, "leafbuild_traverse_bvh", count.value)
    return out, count.value


def leafbuild_traverse_bvh(position, spec, morton=MORTON_MASKED, ctx=None):
    """leafbuild_traverse_bvh(position, spec) (BVHTraverse.jl:1423-1428): pairs
    (i, j, d) with s < r*r, 0-based original indices."""
    out, n = neighborlist_device(position, spec, PRED_SCALAR, morton, ctx)
    return out.to_host(n)


def simdbuild_traverse_bvh(position, spec, morton=MORTON_MASKED, ctx=None):
    """simdbuild_traverse_bvh (ext/NaiveSIMD.jl:706-710): the 4x4-tile path with the
    sqrt(s) < r test.  Like the reference (NaiveSIMD.jl:425) it requires apl == 4."""
    if spec.atomsperleaf != 4:
        raise SpecError(_lib.ND_ERR_ARG, "This simd method requires 4 primitives per leaf")
    out, n = neighborlist_device(position, spec, PRED_SIMD, morton, ctx)
    return out.to_host(n)


def exptsimdbuild_traverse_bvh(position, spec, morton=MORTON_MASKED, ctx=None):
    """exptsimdbuild_traverse_bvh (ext/NaiveSIMD.jl:399-403): SoA result (i, j, d2)
    where d2 holds the distance, as the reference's field does."""
    p = simdbuild_traverse_bvh(position, spec, morton, ctx)
    return {"i": p["i"].copy(), "j": p["j"].copy(), "d2": p["d"].copy()}


def leafTreeData(position, spec, morton=MORTON_MASKED, ctx=None):
    """leafTreeData (BVHTraverse.jl:544-597) -> (keys, (index, morton_code, position)
    sorted, mortsort).  Builds on the device; the traversal result is dropped."""
    ctx = _ctx(ctx)
    neighborlist_device(position, spec, PRED_SCALAR, morton, ctx)
    return tree_export(spec, position, ctx)


def tree_export(spec, position, ctx=None):
    ctx = _ctx(ctx)
    nodes = np.zeros(2 * spec.leaves_count - 1, NODE_DTYPE)
    perm = np.zeros(spec.atom_count, np.int32)
    codes = np.zeros(spec.atom_count, np.int32)
    rc = load().nd_tree_export(ctx.handle, ptr(nodes), ptr(perm), ptr(codes))
    check(rc, ctx, "tree_export")
    x = position.to_host() if isinstance(position, DeviceArray) else np.asarray(position, np.float32)
    prim = {"index": perm.copy(), "morton_code": codes, "position": x.reshape(-1, 3)[perm]}
    return nodes, prim, perm


def _pairs_with_capacity(call, what, ctx, cap):
    """Runs call(out_ptr, cap, count_ref) under the capacity protocol; returns
    (DeviceArray of pairs, count), the result left on the device."""
    out = _pairbuf.get(ctx, cap)
    count = ctypes.c_int64(0)
    rc = call(out.ptr, out.shape[0], ctypes.byref(count))
    if rc == _lib.ND_ERR_CAPACITY:
        out = _pairbuf.get(ctx, count.value + count.value // 8 + 1024)
        rc = call(out.ptr, out.shape[0], ctypes.byref(count))
    check(rc, ctx, what, count.value)
    return out, count.value


def pointquery_device(position, spec, morton=MORTON_MASKED, query=QUERY_OWN_LEAF, ctx=None):
    """Point-query list on device (nd_build_traverse_bvh); returns (DeviceArray
    of pairs, count)."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    if d.shape[0] != spec.atom_count:
        raise SpecError(_lib.ND_ERR_ARG, "position length does not match atom_count")
    L = load()
    return _pairs_with_capacity(
        lambda o, cap, cnt: L.nd_build_traverse_bvh(ctx.handle, d.ptr, ctypes.byref(spec.c), query, morton,
                                                    o, cap, cnt),
        "build_traverse_bvh", ctx, max(1024, 48 * spec.atom_count))


def build_traverse_bvh(position, spec, morton=MORTON_MASKED, query=QUERY_OWN_LEAF, ctx=None):
    """build_traverse_bvh(position, spec) (BVHTraverse.jl:1416-1421): TreeData's
    +-r tree, then neighbor_traverse's point queries from each atom's own leaf
    (query=QUERY_ROOT: the root-started walk of gpuneighbor_traverse!).  Pairs
    (query, subject, d) with 0-based original indices, query-major."""
    out, n = pointquery_device(position, spec, morton, query, ctx)
    return out.to_host(n)


def tree_store(spec, ctx=None):
    """store of the last one-call build: nL-1 int32 CAS slots (1-based leaves)."""
    ctx = _ctx(ctx)
    st = np.zeros(spec.leaves_count - 1, np.int32)
    check(load().nd_tree_store_export(ctx.handle, ptr(st)), ctx, "tree_store")
    return st


def TreeData(position, spec, morton=MORTON_MASKED, ctx=None):
    """TreeData(position, spec) (BVHTraverse.jl:500-543) -> (keys, (index,
    morton_code, position) sorted, mortsort, store).  Built on the device by the
    one-call point-query list, whose pairs are dropped."""
    ctx = _ctx(ctx)
    build_traverse_bvh(position, spec, morton, QUERY_OWN_LEAF, ctx)
    nodes, prim, perm = tree_export(spec, position, ctx)
    return nodes, prim, perm, tree_store(spec, ctx)


def neighbor_traverse(keys, positions, spec, query=QUERY_OWN_LEAF, ctx=None):
    """neighbor_traverse(keys, positions, spec) (BVHTraverse.jl:1125-1234) over
    a given tree; positions is TreeData's sorted primitive dict (index,
    position)."""
    ctx = _ctx(ctx)
    nodes = DeviceArray.from_host(ctx, np.ascontiguousarray(keys, NODE_DTYPE))
    x = DeviceArray.from_host(ctx, np.ascontiguousarray(positions["position"], np.float32).reshape(-1, 3))
    ix = DeviceArray.from_host(ctx, np.ascontiguousarray(positions["index"], np.int32))
    L = load()
    out, n = _pairs_with_capacity(
        lambda o, cap, cnt: L.nd_neighbor_traverse(ctx.handle, nodes.ptr, x.ptr, ix.ptr, ctypes.byref(spec.c),
                                                   query, o, cap, cnt),
        "neighbor_traverse", ctx, max(1024, 48 * spec.atom_count))
    return out.to_host(n)


def quantized_positions(position, ctx=None):
    """quantized_positions! (BVHTraverse.jl:290-314) -> (N, 3) int32: the
    1-based rank of each atom in the stable sorts by x, then y, then z."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    n = d.shape[0]
    q = DeviceArray(ctx, (n, 3), np.int32)
    check(load().nd_quantized_positions(ctx.handle, d.ptr, n, q.ptr), ctx, "quantized_positions!")
    return q.to_host()


def gpubvh_neighborlist(backend, position, spec, morton=MORTON_MASKED, ctx=None):
    """gpubvh_neighborlist(backend, position, spec) (ext/NaiveKA.jl:470-558): returns
    {pairlist, treedata: {tree, mortonpositions, quantizedposition, store}}.
    The pairs are the root-started point queries of gpuneighbor_traverse!
    (:352-452) over the +-r tree (root box kept); quantizedposition is
    quantized_positions! of the input; store holds the build's CAS slots.
    `backend` is accepted for signature parity and must be None or "hip"."""
    if backend not in (None, "hip", "rocm", "gpu"):
        raise SpecError(_lib.ND_ERR_ARG, "only the HIP backend is available")
    ctx = _ctx(ctx)
    quantized = quantized_positions(position, ctx)
    pairs = build_traverse_bvh(position, spec, morton, QUERY_ROOT, ctx)
    nodes, prim, _ = tree_export(spec, position, ctx)
    return {"pairlist": pairs,
            "treedata": {"tree": nodes, "mortonpositions": prim, "quantizedposition": quantized,
                         "store": tree_store(spec, ctx)}}


def threshold_unique_pairs(position, threshold, ctx=None):
    """threshold_pairs(unique_pairs(position), threshold) (AllToAll.jl:65-88):
    every i < j with sqrt(s) < threshold, in (i, j) order.  O(N^2)."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    n = d.shape[0]
    count = ctypes.c_int64(0)
    rc = load().nd_unique_threshold_pairs(ctx.handle, d.ptr, n, np.float32(threshold), None, 0,
                                          ctypes.byref(count))
    if rc == _lib.ND_OK:
        return np.empty(0, PAIR_DTYPE)
    if rc != _lib.ND_ERR_CAPACITY:
        check(rc, ctx, "threshold_pairs")
    out = DeviceArray(ctx, (count.value,), PAIR_DTYPE)
    rc = load().nd_unique_threshold_pairs(ctx.handle, d.ptr, n, np.float32(threshold), out.ptr,
                                          count.value, ctypes.byref(count))
    check(rc, ctx, "threshold_pairs")
    return out.to_host(count.value)


def canonical(pairs):
    """Order-free form of a pair list: rows (min(i,j), max(i,j), bits(d)), sorted."""
    i = np.minimum(pairs["i"], pairs["j"]).astype(np.int64)
    j = np.maximum(pairs["i"], pairs["j"]).astype(np.int64)
    d = np.ascontiguousarray(pairs["d"]).view(np.uint32).astype(np.int64)
    a = np.stack([i, j, d], axis=1)
    return a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))]


__all__ = ["SpheresBVHSpecs", "binwidth", "mortoncodes", "sortperm", "bounding_volume_hierarchy",
           "leafneighbor_traverse", "leafbuild_traverse_bvh", "simdbuild_traverse_bvh",
           "exptsimdbuild_traverse_bvh", "leafTreeData", "gpubvh_neighborlist", "TreeData",
           "build_traverse_bvh", "neighbor_traverse", "quantized_positions", "tree_store",
           "pointquery_device",
           "threshold_unique_pairs", "neighborlist_device", "canonical", "tree_export",
           "CapacityError"]
