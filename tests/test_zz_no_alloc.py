"""GPU: nd_leafbuild_traverse_bvh_noalloc, the one-call neighbour list with
its work buffers in the library's static array (no allocation), against the
stream-ordered form that allocates its own: the same count, status and
pairs in the same order, on two calls in a row (the second smaller, so the
array holds the first call's data when it starts), and a size beyond the
array rejected."""
import ctypes

import numpy as np
import pytest

from conftest import pairs_to_rows

pytestmark = pytest.mark.gpu


def _rand(n, L, seed):
    return (np.random.default_rng(seed).random((n, 3)) * L).astype(np.float32)


def _call(pkg, lib, fn, ctx, x, apl, pred, mode, cap):
    n = len(x)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=n, atomsperleaf=apl)
    d = pkg.DeviceArray.from_host(ctx, x)
    cnt = pkg.DeviceArray.from_host(ctx, np.full(1, -5, np.int64))
    st = pkg.DeviceArray.from_host(ctx, np.full(1, 99, np.int32))
    out = pkg.DeviceArray(ctx, (cap,), pkg.PAIR_DTYPE)
    rc = getattr(lib, fn)(ctx.handle, d.ptr, ctypes.byref(spec.c), pred, mode, out.ptr, cap, cnt.ptr, st.ptr)
    ctx.synchronize()
    if rc != 0:
        return rc, None, None, None
    c = int(cnt.to_host()[0])
    return rc, c, int(st.to_host()[0]), pairs_to_rows(out.to_host(min(c, cap)))


@pytest.mark.parametrize("apl,pred,mode", [(4, 0, 1), (8, 1, 1), (4, 0, 0)])
def test_noalloc_matches_async_twice(pkg, apl, pred, mode):
    ctx = pkg.default_context()
    lib = pkg.library()
    first, second = 60000, 12000     # the second call smaller than the first
    for n, seed in ((first, 31), (second, 32)):
        x = _rand(n, (n / 0.8) ** (1 / 3), seed)
        cap = 80 * n
        ra, ca, sa, pa = _call(pkg, lib, "nd_leafbuild_traverse_bvh_async", ctx, x, apl, pred, mode, cap)
        rn, cn, sn, pn = _call(pkg, lib, "nd_leafbuild_traverse_bvh_noalloc", ctx, x, apl, pred, mode, cap)
        assert ra == 0 and rn == 0
        assert ca == cn and sa == sn == 0 and ca > 0
        assert np.array_equal(pa, pn)
        # the tree export reads the last call's tree (in the static array)
        nodes_n = np.empty(2 * (n // apl) - 1, pkg.NODE_DTYPE)
        perm_n = np.empty(n, np.int32)
        assert lib.nd_tree_export(ctx.handle, nodes_n.ctypes.data_as(ctypes.c_void_p),
                                  perm_n.ctypes.data_as(ctypes.c_void_p), None) == 0
        _call(pkg, lib, "nd_leafbuild_traverse_bvh_async", ctx, x, apl, pred, mode, cap)
        nodes_a = np.empty_like(nodes_n)
        perm_a = np.empty_like(perm_n)
        assert lib.nd_tree_export(ctx.handle, nodes_a.ctypes.data_as(ctypes.c_void_p),
                                  perm_a.ctypes.data_as(ctypes.c_void_p), None) == 0
        assert np.array_equal(nodes_n.view(np.uint8), nodes_a.view(np.uint8))
        assert np.array_equal(perm_n, perm_a)


def test_noalloc_capacity_and_errors(pkg):
    ctx = pkg.default_context()
    lib = pkg.library()
    n = 8000
    x = _rand(n, (n / 0.8) ** (1 / 3), 33)
    # too small an output buffer: the count, nothing written (as the async form)
    rc, c, s, _ = _call(pkg, lib, "nd_leafbuild_traverse_bvh_noalloc", ctx, x, 4, 0, 1, 10)
    ra, ca, _, _ = _call(pkg, lib, "nd_leafbuild_traverse_bvh_async", ctx, x, 4, 0, 1, 80 * n)
    assert rc == 0 and c == ca and s == 0
    # a size whose work buffers exceed the static array is rejected before any work
    big = 1 << 20
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=big, atomsperleaf=4)
    cnt = pkg.DeviceArray.from_host(ctx, np.zeros(1, np.int64))
    st = pkg.DeviceArray.from_host(ctx, np.zeros(1, np.int32))
    d = pkg.DeviceArray(ctx, (big, 3), np.float32)
    rc = lib.nd_leafbuild_traverse_bvh_noalloc(ctx.handle, d.ptr, ctypes.byref(spec.c), 0, 1, None, 0,
                                                cnt.ptr, st.ptr)
    assert rc == pkg._lib.ND_ERR_CAPACITY
    assert b"static scratch" in lib.nd_last_error(ctx.handle)
