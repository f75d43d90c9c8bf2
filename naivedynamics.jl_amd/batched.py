"""Batched ragged replicas: many independent systems advanced by one launch
sequence per step.

The replicas are packed in space, each in its own wall box, with gaps wider
than the neighbour-list radius between boxes.  The packed set is one
simulation: one key sort, one LBVH, one list walk and one fused force +
integrate kernel per step serve every replica, and a rebuild window is one
captured hipGraph.  No listed pair joins two replicas (their atoms stay in
their boxes and the boxes are further apart than the list radius), and each
atom reflects off its own replica's box (nd_sim_set_replicas), so every
replica follows the dynamics it would follow alone.  Uneven sizes cost
nothing extra: the work is balanced per atom, not per replica.

Padding atoms (fewer than one leaf) complete the last LBVH leaf; they sit
far apart on a line past the packed region, carry no charge and no velocity,
and never interact.

The reference has no batched driver (one simulate! call advances one system,
Simulator.jl:154-256); this is the batched form built on the same hot path.
"""
from dataclasses import replace

import numpy as np

from . import _lib, synth
from ._lib import check, load
from .simulator import Simulation


def replica_sizes(nrep, lo=256, hi=16384, seed=11):
    """Sizes drawn log-uniform in [lo, hi] (PCG64 with the given seed)."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size=nrep))).astype(int).clip(lo, hi)


def make_replicas(sizes, seed=11, density=0.8, kT=1.0, charges=False):
    """One system per size at the given density, each with its own seed:
    jittered-lattice positions in [0, L_r)^3 and Gaussian velocities."""
    reps = []
    for r, n in enumerate(sizes):
        n = int(n)
        L = synth.box_side(n, density)
        s = seed * 100003 + r
        reps.append(dict(n=n, L=L, x=synth.lattice_positions(n, L, s),
                         v=synth.gaussian_velocities(n, kT, s),
                         q=synth.balanced_charges(n, s) if charges else np.zeros(n)))
    return reps


def pack(reps, gap, leaf=16):
    """Shelf packing of the replica boxes (largest first) with `gap` between
    boxes.  Returns the packed arrays, per-atom replica ids (-1 for padding),
    the replica boxes (nrep x 6: lo xyz, hi xyz), per-replica atom offsets
    and the global box side."""
    nrep = len(reps)
    side = np.array([r["L"] for r in reps])
    order = np.argsort(-side, kind="stable")
    cell = side + gap
    W = 1.15 * float(np.sum(cell ** 3)) ** (1.0 / 3.0)
    W = max(W, float(cell.max()))
    origin = np.zeros((nrep, 3))
    x0 = y0 = z0 = 0.0
    row_d = layer_h = 0.0
    for r in order:
        c = cell[r]
        if x0 + c > W:                      # next row
            x0 = 0.0
            y0 += row_d
            row_d = 0.0
        if y0 + c > W:                      # next layer
            y0 = 0.0
            z0 += layer_h
            layer_h = 0.0
        origin[r] = (x0 + 0.5 * gap, y0 + 0.5 * gap, z0 + 0.5 * gap)
        x0 += c
        row_d = max(row_d, c)
        layer_h = max(layer_h, c)
    n_atoms = int(sum(r["n"] for r in reps))
    npad = (-n_atoms) % leaf
    top = float(max(origin[r][d] + side[r] for r in range(nrep) for d in range(3)))
    n = n_atoms + npad
    x = np.empty((n, 3))
    v = np.zeros((n, 3))
    q = np.zeros(n)
    rep = np.full(n, -1, np.int32)
    boxes = np.empty((nrep, 6))
    start = np.zeros(nrep + 1, np.int64)
    for r, R in enumerate(reps):
        a, b = start[r], start[r] + R["n"]
        start[r + 1] = b
        x[a:b] = np.asarray(R["x"], np.float64) + origin[r]
        v[a:b] = R["v"]
        q[a:b] = R["q"]
        rep[a:b] = r
        boxes[r, :3] = origin[r]
        boxes[r, 3:] = origin[r] + side[r]
    for k in range(npad):                   # inert atoms, 4 * gap apart on a line
        x[n_atoms + k] = (top + gap * (1 + 4 * k), top + gap, top + gap)
    hi = float(x.max()) + gap
    return dict(x=x, v=v, q=q, rep=rep, boxes=boxes, start=start, n=n, n_atoms=n_atoms,
                npad=npad, L=hi, origin=origin)


class BatchedSimulation:
    """nrep replicas in one nd_sim: set_state from the replicas, run(k) for all,
    replica_states() in each replica's own coordinates."""

    def __init__(self, reps, spec, ctx=None, max_neighbors=0):
        gap = max(spec.rc, spec.rc_coul) + spec.skin + 1.0
        self.reps = reps
        self.packed = pack(reps, gap, spec.atomsperleaf)
        P = self.packed
        self.spec = replace(spec, lo=0.0, hi=P["L"], walls=True)
        self.sim = Simulation(P["n"], self.spec, ctx=ctx, max_neighbors=max_neighbors)
        self.ctx = self.sim.ctx
        rep = np.ascontiguousarray(P["rep"], np.int32)
        boxes = np.ascontiguousarray(P["boxes"], np.float64)
        check(load().nd_sim_set_replicas(self.sim.handle, len(reps), _lib.ptr(rep), _lib.ptr(boxes)),
              self.ctx, "nd_sim_set_replicas")

    def set_state(self):
        P = self.packed
        self.sim.set_state(P["x"], P["v"], P["q"])

    def run(self, nsteps):
        self.sim.run(nsteps)

    def prepare(self):
        self.sim.prepare()

    def energies(self):
        return self.sim.energies()

    def stats(self):
        return self.sim.stats()

    def replica_states(self):
        """[(x, v, f)] per replica, positions relative to the replica's box origin."""
        x, v, f = self.sim.get_state()
        P = self.packed
        out = []
        for r in range(len(self.reps)):
            a, b = P["start"][r], P["start"][r + 1]
            out.append((x[a:b] - P["origin"][r], v[a:b], f[a:b]))
        return out

    def close(self):
        self.sim.close()


def autotune(reps, spec, candidates=None, windows=3, ctx=None, log=None, tiles=True):
    """Pick the packed run's traversal and force tile shapes by timing
    `windows` rebuild windows of each candidate after one warm-up window (every
    window graph captured before timing).  Stage 1: LBVH leaf size (the
    traversal and force tile: 8, 16 or 32 atoms) with the two-phase (0) or rope
    (1) walk, and the cell list (3, no tree).  Stage 2 (`tiles`), from the
    stage-1 winner, one dimension at a time: the two-phase walk's tree records
    per round trip (`walk_front` 8 or 16) and the step kernel's neighbours per
    pipeline batch (`pair_batch` 2, 4 or 8, FP32 only) and its lanes per atom
    (`atom_lanes` 1 or 2).  Returns (best spec, [(leaf, walk, walk_front,
    pair_batch, atom_lanes, ms_per_step)]); 0 in the shape fields means the
    default (8 records, 4 neighbours, 1 lane)."""
    if candidates is None:
        candidates = [(a, w) for a in (8, 16, 32) for w in (0, 1)] + [(16, 3)]
    table = []

    def measure(s2):
        b = BatchedSimulation(reps, s2, ctx=ctx)
        b.set_state()
        W = s2.rebuild_every
        b.run(W)
        b.prepare()                         # every window graph captured before timing
        b.ctx.synchronize()
        b.ctx.timer_start()
        b.run(windows * W)
        ms = b.ctx.timer_stop() / (windows * W)
        b.close()
        row = (s2.atomsperleaf, s2.list_walk, s2.walk_front, s2.pair_batch, s2.atom_lanes, ms)
        table.append(row)
        if log:
            log("autotune: leaf %d walk %d front %d batch %d lanes %d: %.4f ms/step" % row)
        return ms

    best, best_ms = None, None
    for apl, walk in candidates:
        s2 = replace(spec, atomsperleaf=apl, list_walk=walk)
        ms = measure(s2)
        if best is None or ms < best_ms:
            best, best_ms = s2, ms
    if tiles:
        dims = []
        if best.list_walk == 0:
            dims.append(("walk_front", (8, 16)))
        if best.precision == "fp32":
            dims.append(("pair_batch", (2, 4, 8)))
        dims.append(("atom_lanes", (1, 2)))
        defaults = {"walk_front": 8, "pair_batch": 4, "atom_lanes": 1}
        for name, values in dims:
            cur = getattr(best, name) or defaults[name]
            for v in values:
                if v == cur:
                    continue
                s2 = replace(best, **{name: v})
                ms = measure(s2)
                if ms < best_ms:
                    best, best_ms = s2, ms
    return best, table


def run_one_per_replica(reps, spec, nsteps, ctx=None):
    """Baseline: every replica is its own simulation (padded to whole leaves
    the same way), with its own launches per step, advanced one after
    another.  Returns per-replica seconds of device time."""
    times = []
    for R in reps:
        b = BatchedSimulation([R], spec, ctx=ctx)
        b.set_state()
        b.run(spec.rebuild_every)           # warm-up window
        b.prepare()                         # every window graph captured before timing
        b.ctx.synchronize()
        b.ctx.timer_start()
        b.run(nsteps)
        times.append(b.ctx.timer_stop() / 1e3)
        b.close()
    return np.array(times)


__all__ = ["replica_sizes", "make_replicas", "pack", "BatchedSimulation", "autotune",
           "run_one_per_replica"]
