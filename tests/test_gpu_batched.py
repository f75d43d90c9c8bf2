"""GPU: batched ragged replicas packed into one simulation.

Each replica of a packed run must follow the FP64 oracle's trajectory of that
replica alone (walls at its own box), so no pair may cross replicas and every
atom must reflect off its own replica's box.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [256, 300, 517, 1024, 1500, 2048, 257]


def _spec(pkg, precision="fp64", **kw):
    d = dict(duration=0, stepwidth=0.001, rc=2.5, rc_coul=0.0, skin=0.3, rebuild_every=10,
             precision=precision, energy_every=10, atomsperleaf=16)
    d.update(kw)
    return pkg.SimSpec(**d)


@pytest.mark.parametrize("walk", [0, 3])
def test_packed_replicas_follow_the_oracle(pkg, O, walk):
    # walk 0: two-phase LBVH walk; 3: cell list over the packed box
    reps = pkg.batched.make_replicas(SIZES, seed=5)
    # the first layer in x of every replica moves outwards fast: walls are hit
    for R in reps:
        R["v"] = R["v"].copy()
        R["v"][R["x"][:, 0] < 0.8, 0] = -20.0
    b = pkg.BatchedSimulation(reps, _spec(pkg, list_walk=walk))
    b.set_state()
    b.run(40)
    states = b.replica_states()
    e = b.energies()
    pe_sum = 0.0
    for R, (x, v, f) in zip(reps, states):
        xr, vr, fr, E = O.vv_run64(R["x"], R["v"], None, None, O.o64_params(rc_lj=2.5), 0.001, 40,
                                   walls=True, lo=0.0, hi=R["L"])
        assert np.abs(x - xr).max() < 1e-8
        assert np.abs(v - vr).max() < 1e-6
        assert x.min() >= 0.0 and x.max() <= R["L"]
        pe_sum += E[-1, 1]
    # one energy row per 10 steps, summed over all replicas (padding adds nothing)
    assert abs(e[-1, 2] - pe_sum) < 1e-9 * abs(pe_sum)
    b.close()


def test_packed_graph_and_eager_identical(pkg):
    reps = pkg.batched.make_replicas(SIZES[:5], seed=6)
    out = []
    for g in (True, False):
        b = pkg.BatchedSimulation(reps, _spec(pkg, "fp32", use_graph=g))
        b.set_state()
        b.run(7)
        b.run(33)
        out.append(np.concatenate([np.concatenate(s) for s in b.replica_states()]))
        b.close()
    assert np.array_equal(out[0], out[1])


def test_set_replicas_rejects_bad_input(pkg):
    reps = pkg.batched.make_replicas([256, 256], seed=7)
    b = pkg.BatchedSimulation(reps, _spec(pkg))
    rep = np.full(b.packed["n"], 5, np.int32)   # replica index out of range
    with pytest.raises(pkg.SpecError):
        pkg._lib.check(pkg.library().nd_sim_set_replicas(b.sim.handle, 2, pkg._lib.ptr(rep),
                                                         pkg._lib.ptr(b.packed["boxes"])),
                       b.ctx, "nd_sim_set_replicas")
    b.close()


def test_autotune_picks_a_measured_candidate(pkg):
    reps = pkg.batched.make_replicas([512, 1024, 700], seed=8)
    best, table = pkg.autotune(reps, _spec(pkg, "fp32"), candidates=[(8, 0), (16, 0), (16, 1), (16, 3)],
                               windows=1)
    # stage 1: the 4 candidates; stage 2: the other walk front (two-phase
    # winner only), the other two pair batches (FP32), two lanes per atom
    assert 4 + 3 <= len(table) <= 4 + 4
    fastest = min(table, key=lambda t: t[5])
    assert (best.atomsperleaf, best.list_walk, best.walk_front, best.pair_batch,
            best.atom_lanes) == fastest[:5]
