"""GPU: forces and velocity-Verlet trajectories against the FP64 oracle.

Tolerances (relative force error = ||F_gpu - F_64|| / ||F_64|| over all atoms):
  FP32 form: 2e-4 at L ~ 17-69 (FP32 coordinate differences carry ~|x| * 6e-8)
  MIXED form: 1e-5 (FP64 coordinate differences, FP32 pair maths, FP64 sums)
  FP64 form: 1e-10
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(pkg, n, seed, coul=True):
    L = pkg.synth.box_side(n)
    x = pkg.synth.lattice_positions(n, L, seed)
    v = pkg.synth.gaussian_velocities(n, 1.0, seed)
    q = pkg.synth.balanced_charges(n, seed) if coul else None
    return x, v, q, L


def _spec(pkg, L, precision="fp32", steps=10, coul=True, **kw):
    d = dict(duration=steps, stepwidth=0.001, rc=2.5, rc_coul=2.5 if coul else 0.0, skin=0.3,
             rebuild_every=10, lo=0.0, hi=L, precision=precision, energy_every=1)
    d.update(kw)
    return pkg.SimSpec(**d)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-4), ("mixed", 1e-5), ("fp64", 1e-10)])
def test_initial_forces_vs_oracle(pkg, O, precision, tol):
    x, v, q, L = _cfg(pkg, 32768, 1)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision))
    sim.set_state(x, v, q)
    sim.run(0)
    _, _, f = sim.get_state()
    F, U, W, npairs = O.forces64(x, q, O.o64_params(rc_lj=2.5, rc_coul=2.5))
    assert _rel(f, F) < tol, _rel(f, F)
    e = sim.energies()
    assert len(e) == 1 and e[0, 0] == 0
    # MIXED sums FP32 pair terms in FP64: the virial's cancellations leave ~1e-7
    etol = {"fp32": 1e-4, "mixed": 1e-6, "fp64": 1e-12}[precision]
    assert abs(e[0, 2] - U) < etol * abs(U)
    assert abs(e[0, 3] - W) < etol * abs(W)
    # FP32 / MIXED test r^2 < rc^2 in single precision: a pair within an ulp
    # of the cutoff may land on the other side
    assert abs(int(e[0, 4]) - npairs) <= (0 if precision == "fp64" else 2)
    sim.close()


@pytest.mark.parametrize("rc,rcc", [(2.5, 2.0), (2.0, 2.5)])
@pytest.mark.parametrize("precision,tol", [("fp32", 2e-4), ("mixed", 1e-5), ("fp64", 1e-10)])
def test_unequal_cutoffs_vs_oracle(pkg, O, precision, tol, rc, rcc):
    # LJ and Coulomb masked separately (the equal-cutoff case uses one mask)
    x, v, q, L = _cfg(pkg, 16384, 3)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, rc=rc, rc_coul=rcc))
    sim.set_state(x, v, q)
    sim.run(0)
    _, _, f = sim.get_state()
    F, U, W, npairs = O.forces64(x, q, O.o64_params(rc_lj=rc, rc_coul=rcc))
    assert _rel(f, F) < tol, _rel(f, F)
    e = sim.energies()
    etol = {"fp32": 1e-4, "mixed": 1e-6, "fp64": 1e-12}[precision]
    assert abs(e[0, 2] - U) < etol * abs(U)
    assert abs(e[0, 3] - W) < etol * abs(W)
    assert abs(int(e[0, 4]) - npairs) <= (0 if precision == "fp64" else 2)
    sim.close()


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("mixed", 1e-7), ("fp64", 1e-10)])
def test_trajectory_vs_oracle(pkg, O, precision, xtol):
    x, v, q, L = _cfg(pkg, 4096, 2)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, steps=50))
    sim.set_state(x, v, q)
    sim.run(50)
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 50,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < xtol
    assert _rel(fg, fr) < (5e-4 if precision == "fp32" else 1e-5)
    e = sim.energies()
    assert len(e) == 51
    ktol = {"fp32": 1e-4, "mixed": 1e-6, "fp64": 1e-10}[precision]
    assert np.allclose(e[:, 1], E[:, 0], rtol=ktol)
    sim.close()


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("fp64", 1e-10)])
@pytest.mark.parametrize("resort_every,list_walk,key", [(0, 1, 2), (50, 0, 2), (50, 1, 2), (0, 0, 1),
                                                        (50, 0, 1)])
def test_refit_windows_and_walks_vs_oracle(pkg, O, precision, xtol, resort_every, list_walk, key):
    # resort_every = 50 with rebuild_every = 10: rebuilds at steps 10..40 and
    # 60..90 refit the boxes of the tree built at 0 and 50 (atom order and
    # topology kept); the rope walk and the child-box walk must both give
    # every pair inside the list radius
    x, v, q, L = _cfg(pkg, 4096, 7)
    # key 1: interleaved (Morton) cells, 2: Hilbert index of the same cells
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, steps=100, resort_every=resort_every,
                                       list_walk=list_walk, morton=key))
    sim.set_state(x, v, q)
    sim.run(100)
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 100,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < xtol
    assert _rel(fg, fr) < (5e-4 if precision == "fp32" else 1e-10)
    e = sim.energies()
    ktol = {"fp32": 1e-4, "fp64": 1e-10}[precision]
    assert np.allclose(e[:, 1], E[:, 0], rtol=ktol)
    assert sim.stats()["rebuilds"] == 11
    sim.close()


@pytest.mark.parametrize("apl", [8, 16, 32])
def test_refit_window_leaf_sizes_vs_oracle(pkg, O, apl):
    """Refit windows of the two-phase walk (leaf and chunk boxes in one
    kernel, the refit fused with the four-wide records) at every leaf size it
    takes: the FP64 trajectory stays within 1e-10 of the oracle's, which a
    single missing pair would break."""
    x, v, q, L = _cfg(pkg, 4096, 7)
    sim = pkg.Simulation(len(x), _spec(pkg, L, "fp64", steps=60, resort_every=50, list_walk=0,
                                       atomsperleaf=apl))
    sim.set_state(x, v, q)
    sim.run(60)
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 60,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < 1e-10
    assert _rel(fg, fr) < 1e-10
    sim.close()


@pytest.mark.parametrize("energy_every", [0, 10])
def test_walk_block_order_vs_oracle(pkg, O, energy_every):
    """The fused walk places its blocks by the previous walk's work estimates.
    70,016 atoms (2,188 leaves of 32) make 274 blocks of 256 (the last one
    half full): XCD slabs of 35 and 34 blocks, so each order has a full
    first slot and a partial, reversed second one.  The order comes from
    the energy finalize launch (energy_every 10) or from a launch of its
    own (0).  Any permutation gives the same lists, so the FP64 trajectory
    stays within 1e-10 of the oracle's; a block walked twice or skipped
    would break that."""
    n = 70016
    x, v, q, L = _cfg(pkg, n, 11)
    sim = pkg.Simulation(n, _spec(pkg, L, "fp64", steps=30, resort_every=20, list_walk=0,
                                  atomsperleaf=32, energy_every=energy_every))
    sim.set_state(x, v, q)
    sim.run(30)
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 30,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < 1e-10
    assert _rel(fg, fr) < 1e-10
    assert sim.stats()["rebuilds"] == 4
    sim.close()


@pytest.mark.parametrize("pair_batch,walk_front", [(2, 8), (8, 8), (4, 16), (8, 16)])
def test_tile_shapes_vs_oracle(pkg, O, pair_batch, walk_front):
    """The autotuner's tile shapes: neighbours per pipeline batch (FP32 step
    kernel) and tree records per round trip (two-phase walk), each run
    against the FP64 oracle with the default shape's tolerances, and with the
    same list of the initial positions (entries, longest list) as the default."""
    x, v, q, L = _cfg(pkg, 4096, 7)
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 100,
                               walls=True, lo=0.0, hi=L)
    lists = []
    for pb, wf in ((0, 0), (pair_batch, walk_front)):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", steps=100, resort_every=50, pair_batch=pb,
                                           walk_front=wf))
        sim.set_state(x, v, q)
        sim.run(0)  # the list of the initial positions
        st = sim.stats()
        lists.append((st["list_entries"], st["max_neighbors"]))
        sim.run(100)
        xg, vg, fg = sim.get_state()
        assert sim.stats()["rebuilds"] == 11
        assert np.abs(xg - xr).max() < 2e-4
        assert _rel(fg, fr) < 5e-4
        e = sim.energies()
        assert np.allclose(e[:, 1], E[:, 0], rtol=1e-4)
        sim.close()
    assert lists[0] == lists[1]


@pytest.mark.parametrize("precision,xtol,ftol", [("fp32", 2e-4, 5e-4), ("mixed", 1e-7, 1e-5),
                                                 ("fp64", 1e-10, 1e-10)])
def test_two_lanes_per_atom_vs_oracle(pkg, O, precision, xtol, ftol):
    """Step kernel with two lanes per atom (each takes alternate 16-bit list
    rows; half 0 adds half 1's sums): the trajectory against the FP64 oracle
    with each precision's tolerances, refit windows included."""
    x, v, q, L = _cfg(pkg, 4096, 7)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, steps=100, resort_every=50, atom_lanes=2))
    sim.set_state(x, v, q)
    sim.run(100)
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 100,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < xtol
    assert _rel(fg, fr) < ftol
    e = sim.energies()
    ktol = {"fp32": 1e-4, "mixed": 1e-6, "fp64": 1e-10}[precision]
    assert np.allclose(e[:, 1], E[:, 0], rtol=ktol)
    assert np.allclose(e[:, 2], E[:, 1], rtol=ktol, atol=ktol * abs(E[0, 1]))
    sim.close()


def test_two_lanes_per_atom_deterministic(pkg):
    x, v, q, L = _cfg(pkg, 16384, 9)
    out = []
    for _ in range(2):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", steps=30, atom_lanes=2))
        sim.set_state(x, v, q)
        sim.run(30)
        out.append(sim.get_state())
        sim.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_tile_shapes_rejected(pkg):
    with pytest.raises(pkg._lib.SpecError):
        pkg.Simulation(4096, _spec(pkg, 17.0, "fp32", pair_batch=3))
    with pytest.raises(pkg._lib.SpecError):
        pkg.Simulation(4096, _spec(pkg, 17.0, "fp32", walk_front=12))
    with pytest.raises(pkg._lib.SpecError):
        pkg.Simulation(4096, _spec(pkg, 17.0, "fp32", atom_lanes=3))


def _radius_pairs(L, rl2f, seed):
    """Sites 12 apart, each with two 4-atom clusters: atom a0 is cluster A's
    corner towards B along a diagonal (random signs) and b0 cluster B's
    corner towards A, |a0 - b0| a few ulp inside or outside the list radius
    in float32 (both with and without a fused multiply-add); returns the
    positions and the number of boundary pairs inside."""
    rng = np.random.default_rng(seed)
    f = np.float32
    xs, inside = [], 0
    for k, site in enumerate(np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)):
        sg = rng.choice([-1.0, 1.0], 3)
        a0 = (f(6.0) + f(12.0) * site.astype(f) + rng.random(3).astype(f)).astype(f)
        want_in = k % 2 == 0
        t = np.sqrt(float(rl2f) / 3.0)
        b0 = (a0 + (sg * t).astype(f)).astype(f)
        for _ in range(200):  # nudge b0 by ulps until s is on the wanted side in both forms
            d = (a0 - b0).astype(f)
            s1 = f(f(f(d[0] * d[0]) + f(d[1] * d[1])) + f(d[2] * d[2]))
            s2 = f(np.float64(d[2]) * d[2] + np.float64(f(np.float64(d[1]) * d[1] + np.float64(f(d[0] * d[0])))))
            ok = (s1 < rl2f and s2 < rl2f) if want_in else (s1 >= rl2f and s2 >= rl2f)
            if ok:
                break
            ax = k % 3
            b0[ax] = np.nextafter(b0[ax], (a0[ax] if want_in else f(1e9) * sg[ax] + a0[ax]), dtype=f)
        assert ok
        inside += want_in
        offs = np.array([[0.0, 0.0, 0.0], [0.35, 0.1, 0.2], [0.15, 0.4, 0.05], [0.3, 0.3, 0.45]])
        xs.append(a0 - (offs * sg).astype(f))
        xs.append(b0 + (offs * sg).astype(f))
    return np.concatenate(xs).astype(np.float64), inside


@pytest.mark.parametrize("apl", [4, 32])
def test_walk_keeps_pairs_at_the_list_radius(pkg, apl):
    """The two-phase walk prunes nodes by the distance from the atom to the
    node's unpadded box (a sphere test with a rounding slack): a neighbour
    that is the nearest corner of its leaf's box, a few ulp inside the list
    radius along a diagonal, is still listed, and one a few ulp outside is
    not; the rope walk (box test) and the cell list give the same count."""
    L = 100.0
    rl = 2.5 + 0.3
    rlist = np.float32(rl + (L + rl) * 9.5367431640625e-07)
    rl2f = np.float32(rlist * rlist)
    x, inside = _radius_pairs(L, rl2f, 17)
    n = len(x)
    # every other pair: the 6 within each cluster, none between sites
    expect = 2 * (inside + 6 * (n // 4))
    counts = []
    for walk in (0, 1, 3):
        sim = pkg.Simulation(n, _spec(pkg, L, "fp32", coul=False, list_walk=walk, atomsperleaf=apl))
        sim.set_state(x, np.zeros_like(x), None)
        sim.run(0)
        counts.append(sim.stats()["list_entries"])
        sim.close()
    assert counts == [expect] * 3, (counts, expect)


@pytest.mark.parametrize("apl", [2, 4, 8, 16, 32])
def test_walks_give_the_same_list(pkg, apl):
    # 0: two-phase child-box walk (leaf sizes 4..32; others use walk 2),
    # 1: rope walk, 2: one-node child-box walk; over both tree keys (the list
    # is the set of pairs within the list radius, whatever the atom order)
    x, v, q, L = _cfg(pkg, 16384, 8)
    counts = []
    for key in (1, 2):
        for walk in (0, 1, 2):
            sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", list_walk=walk, atomsperleaf=apl,
                                               morton=key))
            sim.set_state(x, v, q)
            sim.run(0)
            st = sim.stats()
            counts.append((st["list_entries"], st["max_neighbors"]))
            sim.close()
    assert len(set(counts)) == 1 and counts[0][0] > 0


@pytest.mark.parametrize("resort_every", [0, 20])
def test_graph_and_eager_identical(pkg, resort_every):
    x, v, q, L = _cfg(pkg, 8192, 3)
    out = []
    for g in (True, False):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", use_graph=g, energy_every=5,
                                           resort_every=resort_every))
        sim.set_state(x, v, q)
        for k in (7, 33, 40):  # windows start and end at different offsets
            sim.run(k)
        out.append(sim.get_state() + (sim.energies(),))
        sim.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_rebuild_period_and_walls(pkg, O):
    x, v, q, L = _cfg(pkg, 4096, 4)
    # the first lattice layers (0.54 from each x wall) move outwards at 25
    # sigma/tau: they reach the wall within the 40 steps and are reflected
    v = v.copy()
    v[x[:, 0] < 0.8, 0] = -25.0
    v[x[:, 0] > L - 0.8, 0] = 25.0
    res = []
    for every in (1, 10, 20):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp64", steps=40, rebuild_every=every,
                                           skin=0.5 if every > 1 else 0.0))
        sim.set_state(x, v, q)
        sim.run(40)
        xg, vg, fg = sim.get_state()
        assert xg.min() >= 0.0 and xg.max() <= L
        res.append(xg)
        sim.close()
    xr, _, _, _ = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 40,
                             walls=True, lo=0.0, hi=L)
    for xg in res:
        assert np.abs(xg - xr).max() < 1e-8


def test_list_overflow_retry(pkg):
    x, v, q, L = _cfg(pkg, 4096, 5)
    sims = []
    for cap in (0, 5):
        # cap = 5 forces the initial list to overflow and the run to resize
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp64", steps=25), max_neighbors=cap)
        sim.set_state(x, v, q)
        sim.run(25)
        sims.append(sim.get_state())
        st = sim.stats()
        assert st["capacity"] >= st["max_neighbors"]
        sim.close()
    for a, b in zip(*sims):
        assert np.abs(a - b).max() < 1e-12


def test_size_limit_is_an_argument_error(pkg):
    # positions are addressed through a buffer resource with a 31-bit byte
    # range: more than 2^26 atoms is refused before anything is allocated
    n = 2 ** 26 + 16
    with pytest.raises(pkg._lib.SpecError):
        pkg.Simulation(n, _spec(pkg, 1000.0, "fp32"))


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("fp64", 1e-10)])
@pytest.mark.parametrize("splits", [(100,), (25, 75), (20, 30, 50)])
def test_thermostat_trajectory_vs_oracle(pkg, O, precision, xtol, splits):
    """rescale_velocity! every 10 steps (Simulator.jl:119-144, 241-243; T from
    v^2) towards T0 = 1.5 with damping 0.5, LJ + Coulomb, against the FP64
    oracle's thermostatted velocity Verlet; runs split at and between
    thermostat steps (a rescale due at a run's last step is applied by the
    next run)."""
    x, v, q, L = _cfg(pkg, 4096, 6)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, thermostat_every=10, temperature=1.5,
                                       vDamp=0.5, energy_every=10))
    sim.set_state(x, v, q)
    for k in splits:
        sim.run(k)
    xg, vg, fg = sim.get_state()
    e = sim.energies()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 100,
                               walls=True, lo=0.0, hi=L, thermostat_every=10, temperature=1.5,
                               gamma=0.5)
    assert np.abs(xg - xr).max() < xtol
    assert np.abs(vg - vr).max() < (2e-3 if precision == "fp32" else 1e-9)
    # the kinetic energy rows at steps 0, 10, ..., 100 (before each rescale)
    ktol = {"fp32": 1e-4, "fp64": 1e-10}[precision]
    assert np.array_equal(e[:, 0], np.arange(0, 101, 10))
    assert np.allclose(e[:, 1], E[::10, 0], rtol=ktol)
    # the thermostat moved the temperature (initial T ~ 1)
    T = 2 * E[:, 0] / (3 * len(x))
    assert abs(T[-1] - 1.5) < abs(T[0] - 1.5)
    sim.close()


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("fp64", 1e-10)])
@pytest.mark.parametrize("key", [0, 2])
def test_config1_plain_small_vs_oracle(pkg, O, precision, xtol, key):
    """The small plain configuration in full: 4,096 atoms at density 0.8, LJ
    only (eps = sigma = 1, rc = 2.5), dt = 0.001, 100 steps with the list
    rebuilt every step at r_list = rc (no skin), 4 atoms per leaf; the
    reference's masked key (0) and the Hilbert key (2)."""
    c = pkg.synth.plain_small(seed=11)
    sim = pkg.Simulation(c["n"], pkg.SimSpec(duration=c["steps"], stepwidth=c["dt"], rc=c["rc"],
                                             rc_coul=c["rc_coul"], skin=c["skin"],
                                             rebuild_every=c["rebuild_every"], lo=0.0, hi=c["L"],
                                             precision=precision, energy_every=1, atomsperleaf=4,
                                             morton=key))
    sim.set_state(c["x"], c["v"], None)
    sim.run(c["steps"])
    xg, vg, fg = sim.get_state()
    xr, vr, fr, E = O.vv_run64(c["x"], c["v"], None, None, O.o64_params(rc_lj=c["rc"]), c["dt"],
                               c["steps"], walls=True, lo=0.0, hi=c["L"])
    assert np.abs(xg - xr).max() < xtol
    assert _rel(fg, fr) < (5e-4 if precision == "fp32" else 1e-10)
    assert sim.stats()["rebuilds"] == c["steps"] + 1
    e = sim.energies()
    assert np.allclose(e[:, 1], E[:, 0], rtol=1e-4 if precision == "fp32" else 1e-10)
    sim.close()


def test_full_size_forces_vs_oracle(pkg, O):
    """N = 262,144 LJ + Coulomb after 20 steps: FP32 and MIXED forces vs FP64."""
    c = pkg.synth.plain_scale(seed=7)
    for precision, tol in (("fp32", 3e-4), ("mixed", 1e-5)):
        sim = pkg.Simulation(c["n"], _spec(pkg, c["L"], precision, energy_every=10))
        sim.set_state(c["x"], c["v"], c["q"])
        sim.run(20)
        xg, vg, fg = sim.get_state()
        F, U, W, npairs = O.forces64(xg, c["q"], O.o64_params(rc_lj=2.5, rc_coul=2.5))
        assert _rel(fg, F) < tol, (precision, _rel(fg, F))
        e = sim.energies()
        assert int(e[-1, 4]) == npairs
        # the uniform liquid fits every workgroup's records in the smaller
        # LDS buffer (2,664 at most by a CPU count): no spill, no switch
        st = sim.stats()
        assert st["table_spills"] == 0 and st["table_records_big"] == 0
        sim.close()


def test_pair_list_force_api(pkg, O):
    x, v, q, L = _cfg(pkg, 4096, 8)
    xf = x.astype(np.float32)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=len(x), atomsperleaf=4)
    pairs = pkg.leafbuild_traverse_bvh(xf, spec, morton=pkg.MORTON_INTERLEAVE)
    f_lj = np.zeros((len(x), 3), np.float32)
    f_c = np.zeros((len(x), 3), np.float32)
    e_lj = pkg.force_lennardjones_(f_lj, pairs, xf, rc=2.5)
    e_c = pkg.force_coulomb_(f_c, pairs, q, xf, rc=2.5)
    f = np.zeros_like(f_lj)
    pkg.sum_forces_(f, f_lj, f_c)
    x64 = xf.astype(np.float64)
    F, U, W, _ = O.forces64(x64, q, O.o64_params(rc_lj=2.5, rc_coul=2.5))
    assert _rel(f, F) < 1e-4
    assert abs(e_lj + e_c - U) < 1e-4 * abs(U)
    Flj, Ulj, _, _ = O.forces64(x64, None, O.o64_params(rc_lj=2.5))
    assert _rel(f_lj, Flj) < 1e-4 and abs(e_lj - Ulj) < 1e-4 * abs(Ulj)


def test_pair_list_force_api_is_deterministic_and_checks_indices(pkg):
    x, v, q, L = _cfg(pkg, 4096, 8)
    xf = x.astype(np.float32)
    spec = pkg.SpheresBVHSpecs(neighbor_distance=2.5, atom_count=len(x), atomsperleaf=4)
    pairs = pkg.leafbuild_traverse_bvh(xf, spec, morton=pkg.MORTON_INTERLEAVE)
    runs = []
    for _ in range(3):
        f = np.zeros((len(x), 3), np.float32)
        e = pkg.force_coulomb_(f, pairs, q, xf, rc=2.5)
        runs.append((f, e))
    # integer (fixed-point) sums: the same bits whatever order the lanes add in
    for f, e in runs[1:]:
        assert np.array_equal(f.view(np.uint32), runs[0][0].view(np.uint32)) and e == runs[0][1]
    bad = np.array(pairs, copy=True)
    bad["j"][17] = len(x)          # one index past the end
    with pytest.raises(pkg.SpecError):
        pkg.force_lennardjones_(np.zeros((len(x), 3), np.float32), bad, xf, rc=2.5)
    bad["j"][17] = -1
    with pytest.raises(pkg.SpecError):
        pkg.force_coulomb_(np.zeros((len(x), 3), np.float32), bad, q, xf, rc=2.5)
    # the context stays usable after the rejected call
    f = np.zeros((len(x), 3), np.float32)
    pkg.force_coulomb_(f, pairs, q, xf, rc=2.5)
    assert np.array_equal(f.view(np.uint32), runs[0][0].view(np.uint32))


@pytest.mark.parametrize("resort_every", [0, 50])
def test_prepared_graphs_give_the_same_trajectory(pkg, resort_every):
    """nd_sim_prepare captures every window graph up front; the run that
    follows replays them and must advance exactly as a run that captures
    lazily (and as an eager run without graphs)."""
    x, v, q, L = _cfg(pkg, 4096, 12)
    out = []
    for mode in ("prepared", "lazy", "eager"):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", resort_every=resort_every, energy_every=10,
                                           use_graph=mode != "eager"))
        sim.set_state(x, v, q)
        sim.run(5)
        if mode == "prepared":
            sim.prepare()
            sim.prepare()          # idempotent
        sim.run(95)
        out.append(sim.get_state())
        assert sim.stats()["rebuilds"] == 11   # steps 0, 10, ..., 100
        # no graph ever replayed buffers that had moved since its capture
        assert sim.stats()["graph_recaptures"] == 0
        sim.close()
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-3), ("fp64", 1e-9)])
@pytest.mark.parametrize("splits", [(60,), (25, 35)])
def test_fast_atoms_force_rebuilds(pkg, O, precision, xtol, splits):
    """Two atoms fast enough to move more than skin/2 = 0.15 between scheduled
    rebuilds (|v| = 30 and 27: ~0.03 per step): the run is redone with the
    list rebuilt before every force it could no longer serve, and the
    trajectory follows the oracle, which evaluates every pair every step (as
    the reference rebuilds its tree every step, Simulator.jl:351-376)."""
    x, v, q, L = _cfg(pkg, 4096, 13)
    v = v.copy()
    v[100] = (30.0, 0.0, 0.0)
    v[2000] = (0.0, -25.0, 10.0)
    sim = pkg.Simulation(len(x), _spec(pkg, L, precision, energy_every=10))
    sim.set_state(x, v, q)
    for k in splits:
        sim.run(k)
    xg, vg, fg = sim.get_state()
    st = sim.stats()
    xr, vr, fr, E = O.vv_run64(x, v, q, None, O.o64_params(rc_lj=2.5, rc_coul=2.5), 0.001, 60,
                               walls=True, lo=0.0, hi=L)
    assert np.abs(xg - xr).max() < xtol, np.abs(xg - xr).max()
    assert st["stale_reruns"] >= 1 and st["forced_rebuilds"] >= 1, st
    assert st["steps"] == 60
    e = sim.energies()
    assert np.array_equal(e[:, 0], np.arange(0, 61, 10))   # one row per energy step
    sim.close()


def test_ordinary_runs_never_rerun(pkg):
    """At kT = 1 no atom comes near skin/2 within 10 steps: no rerun."""
    x, v, q, L = _cfg(pkg, 32768, 14)
    sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32", energy_every=10))
    sim.set_state(x, v, q)
    sim.run(200)
    st = sim.stats()
    assert st["stale_reruns"] == 0 and st["forced_rebuilds"] == 0 and st["rebuilds"] == 21
    sim.close()


def test_simulate_poslog_frames_match_the_state(pkg):
    """poslog frames recorded on the device (ring of 2, drained twice) equal
    the positions get_state returns after the same number of steps."""
    x, v, q, L = _cfg(pkg, 4096, 9)
    spec = _spec(pkg, L, "fp32", steps=40, record_every=10)
    sysobj = pkg.ObjectCollection(position=x.copy(), velocity=v.copy(), charge=q)
    poslog = pkg.simulate_(sysobj, spec, ring=2)
    assert len(poslog) == 5
    assert np.array_equal(poslog[0], x)
    for k in (1, 3):
        sim = pkg.Simulation(len(x), _spec(pkg, L, "fp32"))
        sim.set_state(x, v, q)
        sim.run(10 * k)
        xs, _, _ = sim.get_state()
        sim.close()
        assert np.array_equal(poslog[k], xs)
    assert np.array_equal(poslog[-1], sysobj.position)


@pytest.mark.parametrize("precision,xtol", [("fp32", 2e-4), ("fp64", 1e-10)])
def test_simulate_poslog_frames_vs_oracle(pkg, O, precision, xtol):
    """poslog of simulate! (Simulator.jl:167, 245) every 7 steps over 50 steps
    (frames at 0, 7, ..., 49 and 50; a ring of 3 drained three times) against
    the FP64 oracle's trajectory at the same steps."""
    x, v, q, L = _cfg(pkg, 4096, 10)
    spec = _spec(pkg, L, precision, steps=50, record_every=7)
    sysobj = pkg.ObjectCollection(position=x.copy(), velocity=v.copy(), charge=q)
    poslog = pkg.simulate_(sysobj, spec, ring=3)
    steps = list(range(0, 50, 7)) + [50]
    assert len(poslog) == len(steps)
    p = O.o64_params(rc_lj=2.5, rc_coul=2.5)
    xr, vr = x.copy(), v.copy()
    for k, (a, b) in enumerate(zip(steps[:-1], steps[1:])):
        if k == 0:
            assert np.array_equal(poslog[0], x)
        xr, vr, _, _ = O.vv_run64(xr, vr, q, None, p, 0.001, b - a, walls=True, lo=0.0, hi=L)
        assert np.abs(poslog[k + 1] - xr).max() < xtol, (b, np.abs(poslog[k + 1] - xr).max())
