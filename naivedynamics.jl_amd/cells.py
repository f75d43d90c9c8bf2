"""Uniform cell-list neighbour search: CellListMap's `neighborlist(x, cutoff)`,
the comparison point the reference benchmarks its LBVH against
(assets/BVHBenchSuite.jl:307; also called at src/MDInput.jl:290).

Runs on the GPU through libnaivehip.so (nd_cell_neighborlist).  Pairs are the
same numpy structured array (i, j, d) as the LBVH entry points return, with
0-based original indices, i < j, once per unordered pair.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PAIR_DTYPE, PRED_SCALAR, PRED_SIMD, check, load
from .bvh import _ctx, _dev_xyz, _pairbuf


def neighborlist_device(position, cutoff, predicate=PRED_SIMD, ctx=None):
    """Cell-list pair list on device; returns (DeviceArray of pairs, count)."""
    ctx = _ctx(ctx)
    d, _ = _dev_xyz(ctx, position)
    n = d.shape[0]
    count = ctypes.c_int64(0)
    # the grow-only pair buffer shared with the LBVH entry points: the result
    # stays valid until the next neighbour-list call on this context
    out = _pairbuf.get(ctx, max(1024, 48 * n))
    rc = load().nd_cell_neighborlist(ctx.handle, d.ptr, n, np.float32(cutoff), predicate, out.ptr,
                                     out.shape[0], ctypes.byref(count))
    if rc == _lib.ND_ERR_CAPACITY:
        out = _pairbuf.get(ctx, count.value + count.value // 8 + 1024)
        rc = load().nd_cell_neighborlist(ctx.handle, d.ptr, n, np.float32(cutoff), predicate,
                                         out.ptr, out.shape[0], ctypes.byref(count))
    check(rc, ctx, "neighborlist", count.value)
    return out, count.value


def neighborlist(position, cutoff, predicate=PRED_SIMD, ctx=None):
    """neighborlist(x, cutoff): every pair with distance < cutoff (predicate
    PRED_SIMD, sqrt(s) < cutoff, the brute-force list's test; PRED_SCALAR tests
    s < cutoff^2 like leafneighbor_traverse)."""
    if predicate not in (PRED_SCALAR, PRED_SIMD):
        raise _lib.SpecError(_lib.ND_ERR_ARG, "unknown pair predicate")
    out, n = neighborlist_device(position, cutoff, predicate, ctx)
    return out.to_host(n) if n > 0 else np.empty(0, PAIR_DTYPE)


__all__ = ["neighborlist", "neighborlist_device"]
