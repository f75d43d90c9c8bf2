"""ctypes binding of libnaivehip.so (the C ABI in include/naivedynamics.h).

The product path has no fallback: if the shared library is missing or no
device is present, every call raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnaivehip.so")

ND_OK, ND_ERR_ARG, ND_ERR_DOMAIN, ND_ERR_CAPACITY, ND_ERR_HIP, ND_ERR_STATE = 0, -1, -2, -3, -4, -5
PRED_SCALAR, PRED_SIMD = 0, 1
MORTON_MASKED, MORTON_INTERLEAVE, MORTON_HILBERT = 0, 1, 2
QUERY_OWN_LEAF, QUERY_ROOT = 0, 1
PREC_FP32, PREC_MIXED, PREC_FP64 = 0, 1, 2

NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,)), ("left", "<i4"),
                       ("skip", "<i4")])
PAIR_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("d", "<f4")])


class NDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


class DomainError(NDError, ValueError):
    """Raised where the reference throws InexactError (coordinate out of range)."""


class SpecError(NDError, ValueError):
    """Raised where the reference calls error() on a bad specification."""


class CapacityError(NDError):
    def __init__(self, code, msg, required):
        super().__init__(code, msg)
        self.required = required


class BvhSpec(ctypes.Structure):
    _fields_ = [("neighbor_distance", ctypes.c_float), ("atom_count", ctypes.c_int32),
                ("leaves_count", ctypes.c_int32), ("branches_count", ctypes.c_int32),
                ("atomsperleaf", ctypes.c_int32)]


class ForceParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_float), ("sigma", ctypes.c_float), ("rc", ctypes.c_float),
                ("k_coul", ctypes.c_float), ("rc_coul", ctypes.c_float),
                ("shift_lj", ctypes.c_int32), ("shift_coul", ctypes.c_int32)]


class SimParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("morton_mode", ctypes.c_int32), ("atomsperleaf", ctypes.c_int32),
                ("rebuild_every", ctypes.c_int32), ("walls", ctypes.c_int32),
                ("energy_every", ctypes.c_int32), ("thermostat_every", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("max_neighbors", ctypes.c_int32),
                ("resort_every", ctypes.c_int32), ("list_walk", ctypes.c_int32),
                ("pair_batch", ctypes.c_int32), ("walk_front", ctypes.c_int32),
                ("atom_lanes", ctypes.c_int32),
                ("eps", ctypes.c_double), ("sigma", ctypes.c_double), ("rc", ctypes.c_double),
                ("k_coul", ctypes.c_double), ("rc_coul", ctypes.c_double),
                ("skin", ctypes.c_double), ("dt", ctypes.c_double), ("lo", ctypes.c_double),
                ("hi", ctypes.c_double), ("temperature", ctypes.c_double),
                ("vdamp", ctypes.c_double)]


class AdjParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("precision", ctypes.c_int32), ("nsteps", ctypes.c_int32),
                ("checkpoint_every", ctypes.c_int32), ("rebuild_every", ctypes.c_int32),
                ("nbins", ctypes.c_int32), ("eps", ctypes.c_double), ("sigma", ctypes.c_double),
                ("rc", ctypes.c_double), ("k_coul", ctypes.c_double), ("rc_coul", ctypes.c_double),
                ("skin", ctypes.c_double), ("dt", ctypes.c_double), ("rdf_rmax", ctypes.c_double),
                ("rdf_h", ctypes.c_double), ("volume", ctypes.c_double)]


class SimStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("rebuilds", ctypes.c_int64),
                ("list_entries", ctypes.c_int64), ("max_neighbors", ctypes.c_int32),
                ("capacity", ctypes.c_int32), ("pairs_in_cutoff", ctypes.c_int64),
                ("stale_reruns", ctypes.c_int64), ("forced_rebuilds", ctypes.c_int64),
                ("graph_recaptures", ctypes.c_int64), ("table_spills", ctypes.c_int32),
                ("table_records_big", ctypes.c_int32)]


_P = ctypes.c_void_p
_i32, _i64, _f32, _sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "nd_ctx_create": [_i32, ctypes.POINTER(_P)],
    "nd_ctx_destroy": [_P],
    "nd_last_error": [_P],
    "nd_version": [],
    "nd_malloc": [_P, _sz, ctypes.POINTER(_P)],
    "nd_free": [_P, _P],
    "nd_memcpy_h2d": [_P, _P, _P, _sz],
    "nd_memcpy_d2h": [_P, _P, _P, _sz],
    "nd_memcpy_d2d": [_P, _P, _P, _sz],
    "nd_memset": [_P, _P, _i32, _sz],
    "nd_synchronize": [_P],
    "nd_ctx_stream": [_P, ctypes.POINTER(_P)],
    "nd_timer_start": [_P],
    "nd_timer_stop": [_P, ctypes.POINTER(_f32)],
    "nd_bvh_spec_init": [ctypes.POINTER(BvhSpec), _f32, _i64, _i64],
    "nd_mortoncodes": [_P, _P, _i32, _i32, _P, _P, _P],
    "nd_sortperm": [_P, _P, _i32, _P, _P],
    "nd_bounding_volume_hierarchy": [_P, _P, _P, ctypes.POINTER(BvhSpec), _i32, _P],
    "nd_leafneighbor_traverse": [_P, _P, _P, _P, ctypes.POINTER(BvhSpec), _i32, _P, _i64,
                                 ctypes.POINTER(_i64)],
    "nd_leafbuild_traverse_bvh": [_P, _P, ctypes.POINTER(BvhSpec), _i32, _i32, _P, _i64,
                                  ctypes.POINTER(_i64)],
    "nd_leafbuild_traverse_bvh_async": [_P, _P, ctypes.POINTER(BvhSpec), _i32, _i32, _P, _i64, _P, _P],
    "nd_leafbuild_traverse_bvh_noalloc": [_P, _P, ctypes.POINTER(BvhSpec), _i32, _i32, _P, _i64, _P, _P],
    "nd_leafbuild_traverse_bvh_host": [_P, _P, ctypes.POINTER(BvhSpec), _i32, _i32, _P, _i64,
                                       ctypes.POINTER(_i64)],
    "nd_tree_export": [_P, _P, _P, _P],
    "nd_neighbor_traverse": [_P, _P, _P, _P, ctypes.POINTER(BvhSpec), _i32, _P, _i64,
                             ctypes.POINTER(_i64)],
    "nd_build_traverse_bvh": [_P, _P, ctypes.POINTER(BvhSpec), _i32, _i32, _P, _i64,
                              ctypes.POINTER(_i64)],
    "nd_tree_store_export": [_P, _P],
    "nd_quantized_positions": [_P, _P, _i32, _P],
    "nd_unique_threshold_pairs": [_P, _P, _i32, _f32, _P, _i64, ctypes.POINTER(_i64)],
    "nd_cell_neighborlist": [_P, _P, _i32, _f32, _i32, _P, _i64, ctypes.POINTER(_i64)],
    "nd_cell_neighborlist_host": [_P, _P, _i32, _f32, _i32, _P, _i64, ctypes.POINTER(_i64)],
    "nd_force_lennardjones": [_P, _P, _P, _i64, _P, _i32, ctypes.POINTER(ForceParams),
                              ctypes.POINTER(ctypes.c_double)],
    "nd_force_coulomb": [_P, _P, _P, _i64, _P, _P, _i32, ctypes.POINTER(ForceParams),
                         ctypes.POINTER(ctypes.c_double)],
    "nd_sum_forces": [_P, _P, _P, _P, _i32],
    "nd_sim_create": [_P, ctypes.POINTER(SimParams), ctypes.POINTER(_P)],
    "nd_sim_destroy": [_P],
    "nd_sim_set_state": [_P, _P, _P, _P, _P],
    "nd_sim_get_state": [_P, _P, _P, _P],
    "nd_sim_run": [_P, _i32],
    "nd_sim_prepare": [_P],
    "nd_sim_energies": [_P, _P, _i32, ctypes.POINTER(_i32)],
    "nd_sim_stats_get": [_P, ctypes.POINTER(SimStats)],
    "nd_sim_kernel_timing": [_P, _i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i64)],
    "nd_sim_time_fused_window": [_P, _i32, ctypes.POINTER(ctypes.c_double)],
    "nd_sim_set_replicas": [_P, _i32, _P, _P],
    "nd_sim_record_positions": [_P, _P, _i32],
    "nd_adj_create": [_P, _P, ctypes.POINTER(_P)],
    "nd_adj_destroy": [_P],
    "nd_adj_set_potential": [_P, ctypes.c_double, ctypes.c_double],
    "nd_adj_forward": [_P, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double)],
    "nd_adj_final_state": [_P, _P, _P],
    "nd_adj_backward": [_P, _P, _P, _P, _P],
}
_RESTYPE = {"nd_last_error": ctypes.c_char_p}

_lib = None


def load(path=LIB_PATH):
    """Load libnaivehip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError("libnaivehip.so not built (%s); run __graft_entry__.build()" % path)
        L = ctypes.CDLL(path)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data_as(_P)


def check(rc, ctx=None, what="", required=None):
    if rc == ND_OK:
        return
    msg = what
    if ctx is not None and ctx.handle:
        m = load().nd_last_error(ctx.handle)
        msg = "%s: %s" % (what, m.decode() if m else "")
    if rc == ND_ERR_DOMAIN:
        raise DomainError(rc, msg)
    if rc == ND_ERR_ARG:
        raise SpecError(rc, msg)
    if rc == ND_ERR_CAPACITY:
        raise CapacityError(rc, msg, required)
    raise NDError(rc, msg)


class Context:
    """One HIP stream + grow-only device scratch (nd_ctx).

    Objects that hold device state of this context (simulations, device
    arrays) register with it; closing the context closes them first, and every
    open context is closed at interpreter exit before the HIP runtime goes
    away, so no handle outlives the context it points into."""

    def __init__(self, device=0):
        self.handle = None
        self._dependents = weakref.WeakSet()
        h = _P()
        rc = load().nd_ctx_create(device, ctypes.byref(h))
        check(rc, None, "nd_ctx_create")
        self.handle = h
        _contexts.add(self)

    def register(self, obj):
        """obj.close() must release obj's device state; called before this context closes."""
        self._dependents.add(obj)

    def close(self):
        if self.handle:
            for obj in list(self._dependents):
                try:
                    obj.close()
                except Exception:
                    pass
            load().nd_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(load().nd_synchronize(self.handle), self, "nd_synchronize")

    def timer_start(self):
        check(load().nd_timer_start(self.handle), self, "nd_timer_start")

    def timer_stop(self):
        ms = _f32(0)
        check(load().nd_timer_stop(self.handle, ctypes.byref(ms)), self, "nd_timer_stop")
        return ms.value


class DeviceArray:
    """A device buffer with a numpy dtype/shape, owned by a Context."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _P()
        check(load().nd_malloc(ctx.handle, max(self.nbytes, 1), ctypes.byref(p)), ctx, "nd_malloc")
        self.ptr = p
        ctx.register(self)

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        if a.nbytes:
            check(load().nd_memcpy_h2d(ctx.handle, d.ptr, a.ctypes.data_as(_P), a.nbytes), ctx,
                  "nd_memcpy_h2d")
        return d

    def to_host(self, count=None):
        shape = self.shape if count is None else (count,) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        if out.nbytes:
            check(load().nd_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(_P), self.ptr,
                                       out.nbytes), self.ctx, "nd_memcpy_d2h")
        return out

    def zero(self):
        check(load().nd_memset(self.ctx.handle, self.ptr, 0, self.nbytes), self.ctx, "nd_memset")

    def free(self):
        if self.ptr:
            if self.ctx.handle:
                load().nd_free(self.ctx.handle, self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.free()
        except Exception:
            pass


_default_ctx = None
_contexts = weakref.WeakSet()


@atexit.register
def _close_contexts():
    for c in list(_contexts):
        try:
            c.close()
        except Exception:
            pass


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
