"""Pair-list force entry points of src/Forces.jl:15-74 (textbook physics).

force_lennardjones_(force, pairslist, position) and
force_coulomb_(force, pairslist, charge, position) overwrite `force` (N x 3
float32, host numpy or DeviceArray) like the reference's `!` functions and
return the FP64 pair energy; sum_forces_(force, f1, f2) adds elementwise.
The reference hard-codes eps = -1e10, sigma = 1e-4, k = 1 and broadcasts a
scalar distance (Forces.jl:6-19); here eps, sigma, cutoffs and k are keyword
arguments and the force is the textbook vector form with Newton's third law.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PAIR_DTYPE, DeviceArray, ForceParams, check, load


def _ctx(ctx):
    return ctx if ctx is not None else _lib.default_context()


def _as_dev(ctx, a, dtype, shape=None):
    if isinstance(a, DeviceArray):
        return a
    a = np.ascontiguousarray(a, dtype)
    if shape is not None:
        a = a.reshape(shape)
    return DeviceArray.from_host(ctx, a)


def _pairs_dev(ctx, pairslist):
    if isinstance(pairslist, tuple) and isinstance(pairslist[0], DeviceArray):
        return pairslist  # (DeviceArray, count)
    p = np.ascontiguousarray(pairslist, PAIR_DTYPE)
    if len(p) == 0:
        return None, 0
    return DeviceArray.from_host(ctx, p), len(p)


def _run(kind, force, pairslist, position, charge, params, ctx):
    ctx = _ctx(ctx)
    x = _as_dev(ctx, position, np.float32, (-1, 3))
    n = x.shape[0]
    host_out = not isinstance(force, DeviceArray)
    f = force if not host_out else DeviceArray(ctx, (n, 3), np.float32)
    pd, npairs = _pairs_dev(ctx, pairslist)
    e = ctypes.c_double(0.0)
    if kind == 0:
        rc = load().nd_force_lennardjones(ctx.handle, f.ptr, pd.ptr if pd else None, npairs,
                                          x.ptr, n, ctypes.byref(params), ctypes.byref(e))
        check(rc, ctx, "force_lennardjones!")
    else:
        q = _as_dev(ctx, charge, np.float32)
        rc = load().nd_force_coulomb(ctx.handle, f.ptr, pd.ptr if pd else None, npairs, x.ptr,
                                     q.ptr, n, ctypes.byref(params), ctypes.byref(e))
        check(rc, ctx, "force_coulomb!")
    if host_out:
        force[...] = f.to_host().reshape(np.shape(force))
    return e.value


def force_lennardjones_(force, pairslist, position, *, eps=1.0, sigma=1.0, rc=2.5, shift=False,
                        ctx=None):
    """force_lennardjones!(force, pairslist, position) (Forces.jl:15-45)."""
    p = ForceParams(eps, sigma, rc, 0.0, 0.0, int(shift), 0)
    return _run(0, force, pairslist, position, None, p, ctx)


def force_coulomb_(force, pairslist, charge, position, *, k=1.0, rc=2.5, shift=False, ctx=None):
    """force_coulomb!(force, pairslist, charge) (Forces.jl:56-66); positions give the
    direction of the textbook force."""
    p = ForceParams(0.0, 0.0, 0.0, k, rc, 0, int(shift))
    return _run(1, force, pairslist, position, charge, p, ctx)


def sum_forces_(force, force1, force2, ctx=None):
    """sum_forces!(force, force1, force2) (Forces.jl:68-74)."""
    ctx = _ctx(ctx)
    host_out = not isinstance(force, DeviceArray)
    a = _as_dev(ctx, force1, np.float32, (-1, 3))
    b = _as_dev(ctx, force2, np.float32, (-1, 3))
    n = a.shape[0]
    f = force if not host_out else DeviceArray(ctx, (n, 3), np.float32)
    check(load().nd_sum_forces(ctx.handle, f.ptr, a.ptr, b.ptr, n), ctx, "sum_forces!")
    if host_out:
        force[...] = f.to_host().reshape(np.shape(force))
    return force


__all__ = ["force_lennardjones_", "force_coulomb_", "sum_forces_"]
