"""simulate! / simulate_bvh! (src/Simulator.jl:154-256, 327-379) on the GPU.

Differences from the reference, all deliberate: the step width is a float
(SimSpec.stepwidth is an Int64 in the reference, Simulator.jl:30-48); forces
are textbook LJ + cutoff Coulomb (Forces.jl is not physical); the integrator
is kick-drift-force-kick velocity Verlet (the reference mixes forces of two
different steps, Simulator.jl:206-217); the neighbour list is an LBVH list
with a skin rebuilt every `rebuild_every` steps; reflecting walls follow
boundary_reflect! (Simulator.jl:81-111) and are applied after the drift.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (MORTON_HILBERT, PREC_FP32, PREC_FP64, PREC_MIXED, SimParams, SimStats,
                   check, load)

PRECISIONS = {"fp32": PREC_FP32, "mixed": PREC_MIXED, "fp64": PREC_FP64}


@dataclass
class SimSpec:
    """SimSpec (Simulator.jl:30-49) with a float step width and the physical
    parameters that the reference hard-codes."""
    duration: int
    stepwidth: float = 0.001
    eps: float = 1.0
    sigma: float = 1.0
    rc: float = 2.5
    k_coul: float = 1.0
    rc_coul: float = 0.0
    skin: float = 0.3
    rebuild_every: int = 10
    atomsperleaf: int = 4
    precision: str = "fp32"
    morton: int = MORTON_HILBERT
    walls: bool = True
    lo: float = 0.0
    hi: float = 1.0
    energy_every: int = 10
    thermostat_every: int = 0
    temperature: float = 1.0
    vDamp: float = 1.0
    use_graph: bool = True
    record_every: int = 0
    resort_every: int = 0
    list_walk: int = 0
    pair_batch: int = 0   # step kernel pipeline batch (0 = 4; FP32: 2, 4 or 8)
    walk_front: int = 0   # two-phase walk records per round trip (0 = 8; or 16)
    atom_lanes: int = 0   # step kernel lanes per atom (0 = 1; or 2)


@dataclass
class ObjectCollection:
    """The fields of GenericObjectCollection (MDInput.jl:33-46) the hot path uses."""
    position: np.ndarray
    velocity: np.ndarray
    charge: np.ndarray = None
    mass: np.ndarray = None
    force: np.ndarray = field(default=None)


class Simulation:
    """Owns one nd_sim: device state in LBVH order, neighbour list, step kernels."""

    def __init__(self, n, spec, ctx=None, max_neighbors=0):
        self.ctx = ctx if ctx is not None else _lib.default_context()
        p = SimParams()
        p.n = int(n)
        p.precision = PRECISIONS[spec.precision] if isinstance(spec.precision, str) else int(spec.precision)
        p.morton_mode = int(spec.morton)
        p.atomsperleaf = int(spec.atomsperleaf)
        p.rebuild_every = int(spec.rebuild_every)
        p.walls = int(bool(spec.walls))
        p.energy_every = int(spec.energy_every)
        p.thermostat_every = int(spec.thermostat_every)
        p.use_graph = int(bool(spec.use_graph))
        p.max_neighbors = int(max_neighbors)
        p.resort_every = int(spec.resort_every)
        p.list_walk = int(spec.list_walk)
        p.pair_batch = int(spec.pair_batch)
        p.walk_front = int(spec.walk_front)
        p.atom_lanes = int(spec.atom_lanes)
        p.eps, p.sigma, p.rc = spec.eps, spec.sigma, spec.rc
        p.k_coul, p.rc_coul = spec.k_coul, spec.rc_coul
        p.skin, p.dt = spec.skin, spec.stepwidth
        p.lo, p.hi = spec.lo, spec.hi
        p.temperature, p.vdamp = spec.temperature, spec.vDamp
        self.params = p
        self.n = int(n)
        self.handle = None
        h = ctypes.c_void_p()
        check(load().nd_sim_create(self.ctx.handle, ctypes.byref(p), ctypes.byref(h)), self.ctx,
              "nd_sim_create")
        self.handle = h
        self.ctx.register(self)

    def close(self):
        if self.handle:
            if self.ctx.handle:  # the context closes its simulations before itself
                load().nd_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x, v, q=None, m=None):
        x = np.ascontiguousarray(x, np.float64).reshape(self.n, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(self.n, 3)
        qq = None if q is None else np.ascontiguousarray(q, np.float64).reshape(self.n)
        mm = None if m is None else np.ascontiguousarray(m, np.float64).reshape(self.n)
        P = _lib.ptr
        check(load().nd_sim_set_state(self.handle, P(x), P(v), P(qq), P(mm)), self.ctx,
              "nd_sim_set_state")

    def get_state(self):
        x = np.empty((self.n, 3), np.float64)
        v = np.empty((self.n, 3), np.float64)
        f = np.empty((self.n, 3), np.float64)
        P = _lib.ptr
        check(load().nd_sim_get_state(self.handle, P(x), P(v), P(f)), self.ctx, "nd_sim_get_state")
        return x, v, f

    def run(self, nsteps):
        check(load().nd_sim_run(self.handle, int(nsteps)), self.ctx, "simulate!")

    def prepare(self):
        """Size the list and capture every rebuild-window graph now, so that
        later runs spend no time on capture (nd_sim_prepare)."""
        check(load().nd_sim_prepare(self.handle), self.ctx, "nd_sim_prepare")

    def energies(self, max_rows=1 << 16):
        rows = np.empty((max_rows, 5), np.float64)
        r = ctypes.c_int32(0)
        check(load().nd_sim_energies(self.handle, _lib.ptr(rows), max_rows, ctypes.byref(r)),
              self.ctx, "nd_sim_energies")
        return rows[: r.value]

    def stats(self):
        s = SimStats()
        check(load().nd_sim_stats_get(self.handle, ctypes.byref(s)), self.ctx, "nd_sim_stats")
        return {k: getattr(s, k) for k, _ in SimStats._fields_}

    def time_fused_window(self, nsteps):
        """Advance one rebuild window of nsteps; returns the fused force+integrate
        kernel's average duration in ms (back-to-back launches between HIP events)."""
        ms = ctypes.c_double(0)
        check(load().nd_sim_time_fused_window(self.handle, int(nsteps), ctypes.byref(ms)),
              self.ctx, "nd_sim_time_fused_window")
        return ms.value

    def kernel_timing(self, enable=-1):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        check(load().nd_sim_kernel_timing(self.handle, int(enable), ctypes.byref(ms),
                                          ctypes.byref(n)), self.ctx, "nd_sim_kernel_timing")
        return ms.value, n.value


def simulate_(sys, spec, ctx=None, ring=64):
    """simulate!(sys, spec, clct) (Simulator.jl:154-256): advances `sys` by
    spec.duration steps in place; returns poslog, the positions every
    `record_every` steps (only the start and end if record_every == 0).

    Frames are written on the device in original atom order
    (nd_sim_record_positions) into a ring of up to `ring` frames, which is
    copied to the host in one transfer whenever it fills and at the end."""
    n = len(sys.position)
    sim = Simulation(n, spec, ctx)
    try:
        sim.set_state(sys.position, sys.velocity, sys.charge, sys.mass)
        poslog = [np.array(sys.position, np.float64, copy=True)]
        stride = spec.record_every if spec.record_every > 0 else spec.duration
        nframes = (spec.duration + stride - 1) // stride if stride > 0 else 0
        slots = max(1, min(ring, nframes))
        buf = _lib.DeviceArray(sim.ctx, (slots, n, 3), np.float64)
        used = 0

        def drain():
            if used:
                frames = buf.to_host(used)
                poslog.extend(frames[k] for k in range(used))

        done = 0
        while done < spec.duration:
            k = min(stride, spec.duration - done)
            sim.run(k)
            done += k
            check(load().nd_sim_record_positions(sim.handle, ctypes.c_void_p(buf.ptr.value + used * n * 24), 1),
                  sim.ctx, "nd_sim_record_positions")
            used += 1
            if used == slots:
                drain()
                used = 0
        drain()
        buf.free()
        x, v, f = sim.get_state()
        sys.position[...] = x.astype(np.asarray(sys.position).dtype)
        sys.velocity[...] = v.astype(np.asarray(sys.velocity).dtype)
        sys.force = f
        sys.energies = sim.energies()
        return poslog
    finally:
        sim.close()


simulate_bvh_ = simulate_  # simulate_bvh! (Simulator.jl:327-379): same path, BVH list

__all__ = ["SimSpec", "ObjectCollection", "Simulation", "simulate_", "simulate_bvh_"]
