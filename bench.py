"""Benchmark: steps/s and pair-evals/s of the per-step hot path at N = 262,144.

Workload (one "step" = one velocity-Verlet step of the whole system; the
neighbour list is rebuilt every 10 steps inside the timed region):
262,144 atoms at reduced density 0.8 (L = 68.94), Lennard-Jones eps = sigma = 1
cut at 2.5 plus cutoff Coulomb (q = +-1, k = 1, cut at 2.5), dt = 0.001, list
radius 2.8 (skin 0.3), reflecting walls, FP32 state and pair maths.  Inputs are
synthetic (seeded jittered lattice, Gaussian velocities at kT = 1) and resident
in HBM before the timed region starts.

Prints one JSON line.  Extra fields: the mixed-precision and FP64 forms on the
same workload, the reference-key (masked Morton) form, a variant that re-sorts
every 50 steps and refits the tree in between, the fused force kernel's
roofline, the three added forms (precision on 1,048,576 clustered atoms,
512 batched ragged replicas, the adjoint on 32,768 atoms), and the threaded
CPU baseline (the reference's CPU path, restated in C under oracle/, timed on
a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X vector FP32
FLOPS_PER_PAIR_LJC = 32        # per in-cutoff pair, LJ + Coulomb (+,-,*,/,sqrt,rsqrt,cmp = 1; FMA = 2)
# Algorithmic HBM bytes of one fused FP32 force+integrate launch:
#   8 B per unique listed pair (two 4-byte directed list entries)
#   + 68 B per atom (16 B position read, 16 B velocity read + write,
#     16 B drifted position write, 4 B list length)
BYTES_PER_PAIR = 8
BYTES_PER_ATOM = 68


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# Leaf size of the simulation's internal LBVH (a tuning parameter of this
# implementation; the reference neighbour-list entry points keep their own).
# LBVH leaf size per workload, from sweeps over 8 / 16 / 32 with the Hilbert
# key: 32 on the uniform liquid (ljc-262144), 16 on the dense blobs
SIM_ATOMS_PER_LEAF = 32
# atoms re-sorted by key (and the tree rebuilt) every 200 steps; the list
# rebuilds between (every 10 steps) refit the same tree's boxes.  Sweep at
# config 2 over 8,000 steps (steps/s, first 1,000-step segment / mean): re-sort
# every 100: 26,603 / 25,860, 150: 26,498 / 25,882, 200: 26,778-27,210 /
# 25,910-25,935, 300: 26,747 / 25,728, 500: 26,462-27,022 / 24,772-25,073,
# 1,000: 28,368 / 23,813, 2,000: 28,532 / 22,782, never: 28,466 / 19,449 (the
# refitted tree of the initial lattice order is the best order while the
# lattice melts, and degrades steadily after it).
SIM_RESORT_EVERY = 200
# the blob workload's atoms move faster: re-sort every 100 steps there
# (500 steps, ms/step at 0 / 100 / 500: FP64 0.418 / 0.406 / 0.409, MIXED
# 0.273 / 0.259 / 0.261, FP32 0.243 / 0.231 / 0.233); batched replicas
# every 500 (0.290 / 0.268 / 0.265)
BLOB_RESORT_EVERY = 100
BATCH_RESORT_EVERY = 500
BLOB_ATOMS_PER_LEAF = 16


def make_spec(pkg, c, precision, key, energy_every=10, use_graph=True, resort_every=SIM_RESORT_EVERY):
    return pkg.SimSpec(duration=0, stepwidth=c["dt"], rc=c["rc"], rc_coul=c["rc_coul"],
                       skin=c["skin"], rebuild_every=c["rebuild_every"], lo=0.0, hi=c["L"],
                       precision=precision, morton=key, energy_every=energy_every,
                       use_graph=use_graph, walls=True, atomsperleaf=SIM_ATOMS_PER_LEAF,
                       resort_every=resort_every)


def timed_run(pkg, sim, steps):
    ctx = sim.ctx
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.timer_start()
    sim.run(steps)
    ms = ctx.timer_stop()
    wall = time.perf_counter() - t0
    return ms, wall


def gpu_form(pkg, c, precision, key, warmup, steps):
    sim = pkg.Simulation(c["n"], make_spec(pkg, c, precision, key))
    sim.set_state(c["x"], c["v"], c["q"])
    sim.run(warmup)
    sim.prepare()   # every rebuild-window graph captured before the timed region
    sim.energies()
    ms, wall = timed_run(pkg, sim, steps)
    rows = sim.energies()
    pairs = float(rows[:, 4].mean()) if len(rows) else float("nan")
    return sim, ms, wall, pairs


def copy_calibration(pkg):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of
    this kernel give the counters' scale factors."""
    ctx = pkg.default_context()
    nb = 256 << 20
    a = pkg.DeviceArray(ctx, (nb,), np.uint8)
    b = pkg.DeviceArray(ctx, (nb,), np.uint8)
    a.zero()
    for _ in range(5):
        pkg._lib.check(pkg.library().nd_memcpy_d2d(ctx.handle, b.ptr, a.ptr, nb), ctx, "copy")
    ctx.synchronize()
    log("copied 5 x %d bytes" % nb)


def measured_traffic():
    """HBM bytes per launch of the fused force kernel from the committed
    counter summary (profiles/force_step_ljc262144_pmc.md), or None."""
    path = os.path.join(ROOT, "profiles", "force_step_ljc262144_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def search_comparison(pkg, c):
    """The one-call LBVH neighbour list (interleaved key, 4 atoms per leaf)
    against the cell list (nd_cell_neighborlist) and the point-query list
    (nd_build_traverse_bvh), N = 262,144, r = r_list,
    unique pairs with s < r^2, on the config-2 initial positions (a jittered
    lattice) and on uniform random positions at the same density (the
    reference benchmarks its searches on uniform points); median wall time of
    5 calls each, inputs resident on the device, results left there."""
    r = float(np.float32(c["rc"] + c["skin"]))
    n = len(c["x"])
    L = (n / 0.8) ** (1 / 3)
    ctx = pkg.default_context()
    spec = pkg.SpheresBVHSpecs(neighbor_distance=r, atom_count=n, atomsperleaf=4)

    def timed(fn):
        count = fn()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3, count

    out = {"n": n, "r": r}
    for name, x in (("lattice", np.ascontiguousarray(c["x"], np.float32)),
                    ("uniform", (np.random.default_rng(7).random((n, 3)) * L).astype(np.float32))):
        d = pkg.DeviceArray.from_host(ctx, x)
        ms_l, n_l = timed(lambda: pkg.neighborlist_device(d, spec, pkg.PRED_SCALAR, pkg.MORTON_INTERLEAVE)[1])
        ms_c, n_c = timed(lambda: pkg.cells.neighborlist_device(d, r, pkg.PRED_SCALAR)[1])
        # the point-query traversal (TreeData + neighbor_traverse), same key and leaves
        ms_p, n_p = timed(lambda: pkg.bvh.pointquery_device(d, spec, pkg.MORTON_INTERLEAVE)[1])
        log("search %s: LBVH %.3f ms (%d pairs), cell list %.3f ms (%d pairs), point query %.3f ms (%d pairs)"
            % (name, ms_l, n_l, ms_c, n_c, ms_p, n_p))
        out[name] = {"lbvh_ms": round(ms_l, 3), "lbvh_pairs": int(n_l), "cell_list_ms": round(ms_c, 3),
                     "cell_list_pairs": int(n_c), "point_query_ms": round(ms_p, 3), "point_query_pairs": int(n_p)}
    return out


def precision_form(pkg, steps=500):
    """1,048,576 atoms in 64 blobs: FP32, mixed (FP32 pair maths, FP64 state and
    sums) and plain FP64 from the same state; force error of FP32 and mixed
    against the FP64 GPU forces, energy drift over `steps` steps, and the
    trajectory RMSD of mixed (and FP32) against FP64 after them."""
    c = pkg.synth.precision_blobs()
    n = c["n"]
    out = {"workload": c["workload"], "n": n, "steps": steps}
    res = {}
    for prec in ("fp64", "mixed", "fp32"):
        spec = pkg.SimSpec(duration=0, stepwidth=c["dt"], rc=c["rc"], rc_coul=c["rc_coul"], skin=c["skin"],
                           rebuild_every=c["rebuild_every"], lo=0.0, hi=c["L"], precision=prec,
                           energy_every=10, walls=True, atomsperleaf=BLOB_ATOMS_PER_LEAF,
                           resort_every=BLOB_RESORT_EVERY)
        sim = pkg.Simulation(n, spec)
        sim.set_state(c["x"], c["v"], c["q"])
        sim.run(0)
        sim.prepare()
        _, _, f0 = sim.get_state()
        e0 = sim.energies()
        ms, _ = timed_run(pkg, sim, steps)
        x, _, _ = sim.get_state()
        e = sim.energies()
        W = c["rebuild_every"]
        res_k = float(np.mean([sim.time_fused_window(W) for _ in range(5)]))
        E0 = e0[0, 1] + e0[0, 2]
        E1 = e[-1, 1] + e[-1, 2]
        st = sim.stats()
        res[prec] = dict(f0=f0, x=x, ms=ms, drift=(E1 - E0) / abs(E0), kms=res_k)
        out[prec + "_list"] = {k: int(st[k]) for k in ("rebuilds", "stale_reruns", "forced_rebuilds",
                                                     "max_neighbors", "capacity")}
        out[prec + "_steps_per_s"] = round(steps / (ms / 1e3), 2)
        out[prec + "_force_kernel_ms"] = round(res_k, 5)
        out[prec + "_energy_drift"] = float("%.3e" % res[prec]["drift"])
        sim.close()
        log("precision %s: %.4f ms/step, energy drift %.2e" % (prec, ms / steps, res[prec]["drift"]))
    F = res["fp64"]["f0"]
    fn = np.maximum(np.linalg.norm(F, axis=1), 1.0)
    for prec in ("mixed", "fp32"):
        d = np.linalg.norm(res[prec]["f0"] - F, axis=1)
        out[prec + "_rel_force_err"] = float("%.3e" % (np.linalg.norm(res[prec]["f0"] - F) / np.linalg.norm(F)))
        out[prec + "_max_atom_force_err"] = float("%.3e" % (d / fn).max())
        out[prec + "_rms_atom_force_err"] = float("%.3e" % np.sqrt(np.mean((d / fn) ** 2)))
        out[prec + "_rmsd_vs_fp64"] = float("%.3e" % np.sqrt(np.mean(np.sum((res[prec]["x"] - res["fp64"]["x"]) ** 2, axis=1))))
    out["mixed_over_fp64"] = round(res["fp64"]["ms"] / res["mixed"]["ms"], 3)
    out["mixed_over_fp64_force_kernel"] = round(res["fp64"]["kms"] / res["mixed"]["kms"], 3)
    out["note"] = ("force errors against the FP64 GPU forces at the initial state (per-atom errors "
                   "relative to max(|F_i|, 1)); RMSD against the FP64 trajectory")
    return out


def batched_form(pkg, steps=200):
    """512 replicas of log-uniform size in [256, 16384] (density 0.8, LJ), packed
    into one launch sequence per step, leaf size and walk autotuned; against one
    simulation (one launch sequence) per replica."""
    B = pkg.batched
    sizes = B.replica_sizes(512, seed=11)
    reps = B.make_replicas(sizes, seed=11)
    spec = pkg.SimSpec(duration=0, stepwidth=0.001, rc=2.5, rc_coul=0.0, skin=0.3, rebuild_every=10,
                       precision="fp32", energy_every=10, atomsperleaf=SIM_ATOMS_PER_LEAF,
                       resort_every=BATCH_RESORT_EVERY)
    best, table = pkg.autotune(reps, spec, windows=2, log=log)
    b = pkg.BatchedSimulation(reps, best)
    b.set_state()
    b.run(20)
    b.prepare()
    ms, _ = timed_run(pkg, b.sim, steps)
    b.close()
    sec = B.run_one_per_replica(reps, best, steps)
    atoms = int(sizes.sum())
    out = {"workload": "lj-replicas-512", "replicas": 512, "atoms": atoms, "steps": steps,
           "autotune": [{"leaf": a, "walk": w, "walk_front": f or 8, "pair_batch": p or 4, "atom_lanes": ln or 1,
                         "ms_per_step": round(t, 4)} for a, w, f, p, ln, t in table],
           "chosen": {"leaf": best.atomsperleaf, "walk": best.list_walk, "walk_front": best.walk_front or 8,
                      "pair_batch": best.pair_batch or 4, "atom_lanes": best.atom_lanes or 1},
           "packed_steps_per_s": round(steps / (ms / 1e3), 2),
           "packed_atom_steps_per_s": float("%.4e" % (atoms * steps / (ms / 1e3))),
           "one_per_replica_atom_steps_per_s": float("%.4e" % (atoms * steps / sec.sum())),
           "speedup": round(sec.sum() / (ms / 1e3), 2), "by_size": []}
    for lo, hi in [(256, 512), (512, 1024), (1024, 2048), (2048, 4096), (4096, 8192), (8192, 16385)]:
        m = (sizes >= lo) & (sizes < hi)
        if m.any():
            out["by_size"].append({"size": [lo, hi - 1], "replicas": int(m.sum()),
                                   "one_per_replica_ms_per_step": round(float(np.median(sec[m])) / steps * 1e3, 4),
                                   "one_per_replica_atom_steps_per_s": float("%.3e" % ((sizes[m] * steps).sum() / sec[m].sum())),
                                   "packed_atom_steps_per_s": float("%.3e" % ((sizes[m] * steps).sum() / (ms / 1e3)))})
    log("batched: packed %.4f ms/step, %.1fx one launch sequence per replica" % (ms / steps, out["speedup"]))
    return out


def adjoint_form(pkg, steps=2000):
    """32,768 atoms, LJ + Coulomb (shifted force), `steps` steps, checkpoints
    every 50: loss = squared distance of the final smoothed RDF to the RDF of
    a run with eps = 1.1, sigma = 0.95; gradients by the adjoint against
    central finite differences of the same forward model."""
    n = 32768
    L = pkg.synth.box_side(n)
    x = pkg.synth.lattice_positions(n, L, 5)
    v = pkg.synth.gaussian_velocities(n, 1.0, 5)
    q = pkg.synth.balanced_charges(n, 5)
    out = {"workload": "ljc-adjoint-32768", "n": n, "steps": steps, "checkpoint_every": 50}
    for prec in ("fp32", "fp64"):
        adj = pkg.Adjoint(n, steps, dt=0.001, precision=prec, checkpoint_every=50, rebuild_every=10,
                          nbins=50, rdf_rmax=2.5, rdf_h=0.1, volume=L ** 3)
        adj.set_potential(1.1, 0.95)
        _, target = adj.forward(x, v, q)
        adj.set_potential(1.0, 1.0)
        t0 = time.perf_counter()
        loss, _ = adj.forward(x, v, q, target)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        g = adj.backward()
        tb = time.perf_counter() - t0
        r = {"loss": float("%.6e" % loss), "forward_s": round(tf, 3), "backward_s": round(tb, 3),
             "dL_deps": float("%.6e" % g["deps"]), "dL_dsigma": float("%.6e" % g["dsigma"]),
             "dL_dq77": float("%.6e" % g["dq"][77]), "dL_dx0_norm": float("%.6e" % np.linalg.norm(g["dx0"]))}

        def lossf(eps=1.0, sigma=1.0, qv=q):
            adj.set_potential(eps, sigma)
            return adj.forward(x, v, qv, target)[0]
        fd = {}
        for h in ((1e-3, 1e-7) if prec == "fp64" else (1e-3,)):
            fe = (lossf(eps=1 + h) - lossf(eps=1 - h)) / (2 * h)
            qp, qm = q.copy(), q.copy()
            qp[77] += h
            qm[77] -= h
            fq = (lossf(qv=qp) - lossf(qv=qm)) / (2 * h)
            fd["h=%g" % h] = {"dL_deps": float("%.6e" % fe), "rel_eps": float("%.2e" % (abs(fe - g["deps"]) / abs(fe))),
                              "dL_dq77": float("%.6e" % fq), "rel_q77": float("%.2e" % (abs(fq - g["dq"][77]) / abs(fq)))}
        r["central_fd"] = fd
        out[prec] = r
        adj.close()
        log("adjoint %s: forward %.3f s, backward %.3f s" % (prec, tf, tb))
    out["note"] = ("at 2000 steps the trajectory is chaotic: FD with h = 1e-3 leaves the linear regime; "
                   "FP64 FD with h = 1e-7 is the check of the gradient")
    return out


def adjoint_fd_table(pkg, horizons=(50, 100, 250, 500)):
    """The differentiation form's gradients at its own size (32,768 atoms, LJ +
    Coulomb, shifted force, checkpoints every 50 steps) over growing horizons,
    against central finite differences of the same forward model: FP64 with
    h = 1e-5 and 1e-3, FP32 with h = 1e-3; and the FP32 adjoint against the
    FP64 one.  Relative errors |ad - fd| / |fd| per parameter."""
    n = 32768
    L = pkg.synth.box_side(n)
    x = pkg.synth.lattice_positions(n, L, 5)
    v = pkg.synth.gaussian_velocities(n, 1.0, 5)
    q = pkg.synth.balanced_charges(n, 5)
    names = ["eps", "sigma", "q77", "q12345", "x0[0].x", "x0[4321].z"]
    rows = []
    for T in horizons:
        grads = {}
        for prec, hs in (("fp64", (1e-5, 1e-3)), ("fp32", (1e-3,))):
            adj = pkg.Adjoint(n, T, dt=0.001, precision=prec, checkpoint_every=50, rebuild_every=10,
                              nbins=50, rdf_rmax=2.5, rdf_h=0.1, volume=L ** 3)
            adj.set_potential(1.1, 0.95)
            _, target = adj.forward(x, v, q)
            adj.set_potential(1.0, 1.0)
            adj.forward(x, v, q, target)
            g = adj.backward()
            ad = [g["deps"], g["dsigma"], g["dq"][77], g["dq"][12345], g["dx0"][0, 0], g["dx0"][4321, 2]]
            grads[prec] = ad

            def loss(eps=1.0, sigma=1.0, qv=q, xv=x):
                adj.set_potential(eps, sigma)
                return adj.forward(xv, v, qv, target)[0]

            for h in hs:
                fd = []
                fd.append((loss(eps=1 + h) - loss(eps=1 - h)) / (2 * h))
                fd.append((loss(sigma=1 + h) - loss(sigma=1 - h)) / (2 * h))
                for i in (77, 12345):
                    qp, qm = q.copy(), q.copy()
                    qp[i] += h
                    qm[i] -= h
                    fd.append((loss(qv=qp) - loss(qv=qm)) / (2 * h))
                for i, c in ((0, 0), (4321, 2)):
                    xp, xm = x.copy(), x.copy()
                    xp[i, c] += h
                    xm[i, c] -= h
                    fd.append((loss(xv=xp) - loss(xv=xm)) / (2 * h))
                rel = [abs(a - f) / max(abs(f), 1e-300) for a, f in zip(ad, fd)]
                rows.append({"steps": T, "precision": prec, "h": h, "max_rel_err": float("%.3e" % max(rel)),
                             "rel_err": {k: float("%.2e" % r) for k, r in zip(names, rel)},
                             "adjoint": {k: float("%.6e" % a) for k, a in zip(names, ad)},
                             "fd": {k: float("%.6e" % f) for k, f in zip(names, fd)}})
                log("adjoint fd: %d steps %s h=%g: max rel err %.2e" % (T, prec, h, max(rel)))
            adj.close()
        d = [abs(a - b) / max(abs(b), 1e-300) for a, b in zip(grads["fp32"], grads["fp64"])]
        rows.append({"steps": T, "fp32_vs_fp64_rel": {k: float("%.2e" % r) for k, r in zip(names, d)},
                     "max_rel": float("%.3e" % max(d))})
        log("adjoint fd: %d steps fp32 vs fp64: max rel %.2e" % (T, max(d)))
    return {"workload": "ljc-adjoint-32768", "n": n, "checkpoint_every": 50, "rows": rows}


def cpu_threads():
    """Threads for the CPU leg: the cores this process may run on (its
    affinity mask), capped by OMP_NUM_THREADS where that is set (the host's
    share of cores for one GPU)."""
    aff = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, omp) if omp > 0 else aff), aff, os.cpu_count()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(c, threads, full=False):
    """The reference's threaded CPU path restated in C (oracle/), timed as
    planned for the GPU comparison: per variant a warm-up run, then the median
    of 3 runs of a 50-step window holding 5 list builds (the workload's 1 build
    per 10 steps).  Variants: the interleaved key (the ordering the GPU path
    also sorts by) and the reference's masked key, each with the scalar pair
    test and the 4x4 SIMD tile (sqrt(s) < r).  With the masked key a list
    build takes seconds (its codes are not a locality-preserving order at
    this box size), so outside full mode it is timed once over a 10-step
    window with 1 build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU restatement of the reference path
    O.lib()

    def run(morton, pred, steps):
        bp = O.BaseParams(1.0, 1.0, c["rc"], 1.0, c["rc_coul"], c["dt"], c["skin"], 0.0, c["L"], 4,
                          c["rebuild_every"], 1, pred, morton, threads)
        t0 = time.perf_counter()
        _, _, times, pe = O.base_run(c["x"], c["v"], c["q"], bp, steps)
        return time.perf_counter() - t0, times, pe

    window = 5 * c["rebuild_every"]
    variants = {}
    for key, morton in (("interleaved", O.MORTON_INTERLEAVE), ("masked", O.MORTON_MASKED)):
        for pname, pred in (("simd_tile", O.PRED_SIMD), ("scalar", O.PRED_SCALAR)):
            quick = key == "masked" and not full
            steps = c["rebuild_every"] if quick else window
            if not quick:
                run(morton, pred, c["rebuild_every"])  # warm-up
            runs = [run(morton, pred, steps) for _ in range(1 if quick else 3)]
            runs.sort(key=lambda r: r[0])
            wall, times, pe = runs[len(runs) // 2]
            builds = (steps + c["rebuild_every"] - 1) // c["rebuild_every"]
            variants["%s_%s" % (key, pname)] = {
                "steps_per_s": round(steps / wall, 4), "steps": steps, "list_builds": builds,
                "runs": len(runs), "wall_s": [round(r[0], 3) for r in runs],
                "list_build_s": round((times[0] + times[1] + times[2]) / builds, 4),
                "forces_integration_s_per_step": round(times[3] / steps, 5),
                "pair_evals_per_s": round(pe / wall, 1)}
            log("cpu %s %s: %.4f steps/s (median of %d, %d steps)" % (key, pname, steps / wall, len(runs),
                                                                      steps))
    best = max(("interleaved_simd_tile", "interleaved_scalar"), key=lambda k: variants[k]["steps_per_s"])
    return best, variants


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed/FP64/masked forms")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the masked-key CPU variants with the full protocol too")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the FP32 timed region (for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 5 x 256 MiB device-to-device (counter calibration run)")
    ap.add_argument("--form", choices=("precision", "batched", "adjoint", "adjoint_fd", "search"),
                    help="only this form, with --form-steps steps (for profiler runs)")
    ap.add_argument("--form-steps", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    pkg = ge.load_package()
    if args.copy_calibration:
        copy_calibration(pkg)
        return
    if args.form:
        fn = {"precision": precision_form, "batched": batched_form, "adjoint": adjoint_form,
              "adjoint_fd": lambda pkg: adjoint_fd_table(pkg),
              "search": lambda pkg: search_comparison(pkg, pkg.synth.plain_scale(seed=2))}[args.form]
        out = fn(pkg, args.form_steps) if args.form_steps > 0 else fn(pkg)
        print(json.dumps({"form": args.form, "result": out}, default=float), flush=True)
        return
    c = pkg.synth.plain_scale(seed=2)
    n = c["n"]
    log("workload ljc-262144: N=%d L=%.4f" % (n, c["L"]))

    # ---- headline: FP32 plain path, interleaved LBVH key ----
    sim, ms, wall, pairs = gpu_form(pkg, c, "fp32", pkg.MORTON_HILBERT, args.warmup, args.steps)
    steps_per_s = args.steps / (ms / 1e3)
    log("fp32: %.3f ms/step (%.1f steps/s), wall %.3f s, %.0f pairs in cutoff" %
        (ms / args.steps, steps_per_s, wall, pairs))
    st = sim.stats()
    result = {
        "metric": "steps/s + pair-evals/s at N=262k LJ; mixed vs plain FP64 @1e-5 force err",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ljc-262144", "n": n, "density": 0.8, "rc_lj": c["rc"],
                   "rc_coul": c["rc_coul"], "skin": c["skin"], "dt": c["dt"],
                   "rebuild_every": c["rebuild_every"], "lbvh_key": "hilbert30",
                   "atoms_per_leaf": SIM_ATOMS_PER_LEAF, "resort_every": SIM_RESORT_EVERY, "precision": "fp32",
                   "walls": True},
        "pair_evals_per_s": round(pairs * steps_per_s, 1),
        "pairs_in_cutoff_per_step": pairs,
        "list_pairs": st["list_entries"] // 2,
    }
    if args.profile_only:
        print(json.dumps(result), flush=True)
        return

    # ---- roofline of the fused force + integrate kernel: 20 rebuild windows,
    # each timing 9 back-to-back fused launches between two HIP events ----
    W = c["rebuild_every"]
    times = [sim.time_fused_window(W) for _ in range(20)]
    avg_ms = float(np.mean(times))
    launches = 20 * (W - 1)
    st = sim.stats()
    plist = st["list_entries"] / 2.0
    bytes_launch = BYTES_PER_PAIR * plist + BYTES_PER_ATOM * n
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    flops_launch = FLOPS_PER_PAIR_LJC * pairs
    result["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                          "traffic": measured_traffic(), "kernel": "k_step<float, float, float, 1, false, 4, 1>",
                          "avg_launch_ms": round(avg_ms, 5), "launches": launches,
                          "algorithmic_bytes_per_launch": int(bytes_launch),
                          "fp32_tflops": round(flops_launch / (avg_ms * 1e-3) / 1e12, 2),
                          "fp32_frac": round(flops_launch / (avg_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)}
    log("force kernel: %.4f ms avg over %d launches, %.0f GB/s algorithmic" %
        (avg_ms, launches, achieved))
    sim.close()

    # ---- precision forms and the reference-key form ----
    if not args.no_extra:
        extra = {}
        k = min(args.steps, 300)
        for prec in ("mixed", "fp64"):
            s2, ms2, _, _ = gpu_form(pkg, c, prec, pkg.MORTON_HILBERT, 20, k)
            extra[prec + "_steps_per_s"] = round(k / (ms2 / 1e3), 3)
            W = c["rebuild_every"]
            extra[prec + "_force_kernel_ms"] = round(float(np.mean([s2.time_fused_window(W) for _ in range(10)])), 5)
            s2.close()
            log("%s: %.3f ms/step, force kernel %.4f ms" % (prec, ms2 / k, extra[prec + "_force_kernel_ms"]))
        extra["mixed_over_fp64"] = round(extra["mixed_steps_per_s"] / extra["fp64_steps_per_s"], 3)
        # the pair-force + integrate kernel alone (the list build is shared)
        extra["mixed_over_fp64_force_kernel"] = round(extra["fp64_force_kernel_ms"] / extra["mixed_force_kernel_ms"], 3)
        s2, ms2, _, _ = gpu_form(pkg, c, "fp32", pkg.MORTON_MASKED, 10, 30)
        extra["fp32_masked_key_steps_per_s"] = round(30 / (ms2 / 1e3), 3)
        s2.close()
        log("fp32 masked key: %.3f ms/step" % (ms2 / 30))
        # a full re-sort and tree build at every list rebuild (no refits)
        spec = make_spec(pkg, c, "fp32", pkg.MORTON_HILBERT, resort_every=0)
        s2 = pkg.Simulation(c["n"], spec)
        s2.set_state(c["x"], c["v"], c["q"])
        s2.run(args.warmup)
        s2.prepare()
        ms2, _ = timed_run(pkg, s2, args.steps)
        extra["fp32_resort_every_rebuild_steps_per_s"] = round(args.steps / (ms2 / 1e3), 3)
        s2.close()
        log("fp32 re-sort at every rebuild: %.4f ms/step" % (ms2 / args.steps))
        # the cell-list comparison point: the same steps with lists from a
        # uniform cell list instead of the LBVH walk
        spec = make_spec(pkg, c, "fp32", pkg.MORTON_HILBERT)
        spec.list_walk = 3
        s2 = pkg.Simulation(c["n"], spec)
        s2.set_state(c["x"], c["v"], c["q"])
        s2.run(args.warmup)
        s2.prepare()
        ms2, _ = timed_run(pkg, s2, args.steps)
        extra["fp32_cell_list_steps_per_s"] = round(args.steps / (ms2 / 1e3), 3)
        s2.close()
        log("fp32 cell-list walk: %.4f ms/step" % (ms2 / args.steps))
        extra["search"] = search_comparison(pkg, c)
        result["forms"] = extra
        result["precision_form"] = precision_form(pkg)
        result["batched_form"] = batched_form(pkg)
        result["adjoint_form"] = adjoint_form(pkg)

    # ---- CPU baseline: the reference's threaded CPU path (C restatement) ----
    if not args.no_cpu:
        threads, aff, ncpu = cpu_threads()
        if args.cpu_threads:
            threads = args.cpu_threads
        best, variants = cpu_baseline(c, threads, args.cpu_full)
        b = variants[best]
        m = variants["masked_simd_tile"]
        result["cpu_baseline"] = {
            "value": b["steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "ljc-262144 on %d threads (%s; affinity %d of %d cpus): the reference's threaded CPU "
                      "path restated in C, %s; median of %d runs of %d steps with %d list builds after a "
                      "warm-up run" % (threads, cpu_model(), aff, ncpu, best.replace("_", " "), b["runs"],
                                       b["steps"], b["list_builds"]),
            "pair_evals_per_s": b["pair_evals_per_s"], "list_build_s": b["list_build_s"],
            "reference_key_steps_per_s": m["steps_per_s"], "variants": variants}
        result["gpu_over_cpu"] = round(steps_per_s / b["steps_per_s"], 1)
        mk = result.get("forms", {}).get("fp32_masked_key_steps_per_s")
        if mk:
            result["gpu_over_cpu_reference_key"] = round(mk / m["steps_per_s"], 1)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
