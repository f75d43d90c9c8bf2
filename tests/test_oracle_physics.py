"""CPU: pins the FP64 physics oracle (oracle/oracle64.c) with closed forms,
finite differences, Newton's third law, an O(N^2) numpy evaluation and
velocity-Verlet invariants.  The reference's own force code is not physical
(src/Forces.jl:6-19, 46-53), so nothing from it can pin these values."""
import numpy as np
import pytest


def test_lj_minimum_and_well_depth(O):
    p = O.o64_params(eps=1.3, sigma=0.9, rc_lj=2.5)
    rmin = 2 ** (1 / 6) * 0.9
    x = np.array([[0.0, 0.0, 0.0], [rmin, 0.0, 0.0]])
    F, U, W, n = O.forces64(x, None, p)
    assert n == 1
    assert abs(U + 1.3) < 1e-12          # U(r_min) = -eps
    assert np.abs(F).max() < 1e-12        # F(r_min) = 0
    x[1, 0] = 0.9                         # U(sigma) = 0, F = 24 eps / sigma repulsive
    F, U, W, _ = O.forces64(x, None, p)
    assert abs(U) < 1e-12
    assert abs(F[0, 0] + 24 * 1.3 / 0.9) < 1e-10 and abs(F[1, 0] - 24 * 1.3 / 0.9) < 1e-10


def test_coulomb_pair_closed_form(O):
    p = O.o64_params(rc_lj=0.0, rc_coul=2.5, k_coul=1.0)
    x = np.array([[0.0, 0.0, 0.0], [0.0, 1.7, 0.0]])
    F, U, W, _ = O.forces64(x, np.array([1.0, -1.0]), p)
    assert abs(U + 1 / 1.7) < 1e-14
    assert abs(F[0, 1] - 1 / 1.7 ** 2) < 1e-14   # opposite charges attract
    assert abs(F[1, 1] + 1 / 1.7 ** 2) < 1e-14
    assert abs(W - (-1.7) * (1 / 1.7 ** 2)) < 1e-14  # r_ij . F_ij with r_ij = x_0 - x_1
    x[1, 1] = 2.6                                  # beyond the cutoff
    F, U, W, n = O.forces64(x, np.array([1.0, -1.0]), p)
    assert n == 0 and U == 0.0


def _numpy_forces(x, q, eps, sigma, rc, rcc):
    N = len(x)
    d = x[:, None, :] - x[None, :, :]
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    r = np.sqrt(r2)
    sr6 = (sigma ** 2 / r2) ** 3
    flj = np.where(r < rc, 24 * eps * (2 * sr6 ** 2 - sr6) / r2, 0.0)
    ulj = np.where(r < rc, 4 * eps * (sr6 ** 2 - sr6), 0.0)
    qq = np.outer(q, q) if q is not None else np.zeros((N, N))
    fc = np.where(r < rcc, qq / (r2 * r), 0.0)
    uc = np.where(r < rcc, qq / r, 0.0)
    fs = flj + fc
    F = (fs[:, :, None] * d).sum(1)
    U = 0.5 * (ulj + uc).sum()
    W = 0.5 * np.where(np.isfinite(r2), fs * r2, 0.0).sum()
    return F, U, W


def _liquid(n, seed):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    L = (n / 0.8) ** (1 / 3)
    a = L / side
    sites = rng.choice(side ** 3, n, replace=False)
    iz, rem = np.divmod(sites, side * side)
    iy, ix = np.divmod(rem, side)
    x = (np.stack([ix, iy, iz], 1) + 0.5) * a + rng.uniform(-0.1, 0.1, (n, 3))
    q = rng.permutation(np.r_[np.ones(n // 2), -np.ones(n - n // 2)])
    return x, q, L


def test_cell_list_matches_numpy_all_pairs(O):
    x, q, _ = _liquid(700, 1)
    p = O.o64_params(eps=1.0, sigma=1.0, rc_lj=2.5, rc_coul=2.2, k_coul=0.7)
    F, U, W, _ = O.forces64(x, q, p)
    Fn, Un, Wn = _numpy_forces(x, 0.7 ** 0.5 * q, 1.0, 1.0, 2.5, 2.2)
    assert np.allclose(F, Fn, rtol=1e-11, atol=1e-11)
    assert abs(U - Un) < 1e-9 * abs(Un) and abs(W - Wn) < 1e-9 * abs(Wn)


def test_newton_third_law_and_pair_list_form(O):
    x, q, _ = _liquid(512, 2)
    p = O.o64_params(rc_lj=2.5, rc_coul=2.5)
    F, U, W, n = O.forces64(x, q, p)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()
    i, j = np.triu_indices(len(x), 1)
    F2, U2, W2 = O.forces64_pairs(x, q, i, j, p)
    assert np.allclose(F, F2, rtol=1e-12, atol=1e-12) and abs(U - U2) < 1e-10


def test_force_is_minus_gradient(O):
    x, q, _ = _liquid(64, 3)
    p = O.o64_params(rc_lj=2.5, rc_coul=2.5, shift_lj=True, shift_coul=True)
    F, _, _, _ = O.forces64(x, q, p)
    h = 1e-6
    for (a, d) in [(0, 0), (5, 1), (17, 2), (40, 0)]:
        xp = x.copy(); xp[a, d] += h
        xm = x.copy(); xm[a, d] -= h
        up = O.forces64(xp, q, p)[1]
        um = O.forces64(xm, q, p)[1]
        assert abs(-(up - um) / (2 * h) - F[a, d]) < 1e-5 * max(1.0, abs(F[a, d]))


def test_vv_energy_conservation_and_reversibility(O):
    x, q, L = _liquid(343, 4)
    rng = np.random.default_rng(4)
    v = rng.normal(size=x.shape)
    v -= v.mean(0)
    p = O.o64_params(rc_lj=2.5, rc_coul=0.0, shift_lj=True)
    m = np.ones(len(x))
    x1, v1, F1, E = O.vv_run64(x, v, None, m, p, 0.001, 200)
    tot = E[:, 0] + E[:, 1]
    assert np.abs(tot - tot[0]).max() < 2e-4 * np.abs(E[:, 0]).mean()
    # time reversal: flip velocities and integrate back
    x2, v2, _, _ = O.vv_run64(x1, -v1, None, m, p, 0.001, 200)
    assert np.abs(x2 - x).max() < 1e-8


def test_walls_reflect(O):
    p = O.o64_params(rc_lj=2.5)
    x = np.array([[0.01, 5.0, 5.0], [9.99, 1.0, 1.0]])
    v = np.array([[-20.0, 0.0, 0.0], [30.0, 0.0, 0.0]])
    x1, v1, _, _ = O.vv_run64(x, v, None, np.ones(2), p, 0.001, 1, walls=True, lo=0.0, hi=10.0)
    assert x1[0, 0] == 0.0 and v1[0, 0] == 20.0
    assert x1[1, 0] == 10.0 and v1[1, 0] == -30.0


def test_cpu_baseline_runs_and_conserves(O):
    x, q, L = _liquid(1000, 5)
    v = np.random.default_rng(5).normal(size=x.shape)
    bp = O.BaseParams(1.0, 1.0, 2.5, 1.0, 2.5, 0.001, 0.3, 0.0, L, 4, 10, 1, 0, 1, 4)
    xb, vb, times, pe = O.base_run(x, v, q, bp, 20)
    p = O.o64_params(rc_lj=2.5, rc_coul=2.5)
    xr, vr, _, _ = O.vv_run64(x, v, q, None, p, 0.001, 20, walls=True, lo=0.0, hi=L)
    assert np.abs(xb - xr).max() < 1e-3
    assert pe > 0 and np.all(times >= 0)


def _free_gas(n, seed, L=1000.0):
    """n atoms far apart (no pair inside the cutoff): forces vanish, so the
    only change of velocity is the thermostat's rescaling."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.unravel_index(np.arange(n), (side, side, side)), 1).astype(float)
    x = (g + 0.5) * (L / side)
    v = rng.normal(size=(n, 3))
    return x, v


@pytest.mark.parametrize("gamma", [1.0, 0.5, 0.0])
def test_thermostat_relaxes_temperature_in_closed_form(O, gamma):
    """rescale_velocity! (Simulator.jl:119-144) with T = 2K/(3N): with no
    forces the velocities change only by the rescale, v <- beta v with
    beta^2 = 1 + gamma (T0/T - 1), so after a rescale T' = T + gamma (T0 - T).
    It acts on v_s at every step s > 0 that is a multiple of the period,
    before the kick-drift of step s + 1 (never at s = 0 or within the last
    step of a run)."""
    n, T0 = 64, 2.5
    x, v = _free_gas(n, 6)
    p = O.o64_params(rc_lj=2.5)
    x1, v1, F, E = O.vv_run64(x, v, None, None, p, 0.001, 25, thermostat_every=10,
                              temperature=T0, gamma=gamma)
    assert np.abs(F).max() == 0.0
    T = 2.0 * E[:, 0] / (3 * n)
    T_init = T[0]
    # steps 0..10 unchanged (no rescale at step 0), 11..20 after one rescale
    assert np.allclose(T[:11], T_init, rtol=1e-13)
    T1 = T_init + gamma * (T0 - T_init)
    assert np.allclose(T[11:21], T1, rtol=1e-12)
    T2 = T1 + gamma * (T0 - T1)
    assert np.allclose(T[21:], T2, rtol=1e-12)
    # the direction of every velocity is kept (a uniform scale)
    beta = np.sqrt(T2 / T_init)
    assert np.allclose(v1, beta * v, rtol=1e-12)
    # the rescale at the run's last step (25 is no multiple; 20 is) was applied
    # before step 21; a run of exactly 20 steps ends before it
    _, v20, _, E20 = O.vv_run64(x, v, None, None, p, 0.001, 20, thermostat_every=10,
                                temperature=T0, gamma=gamma)
    assert np.allclose(2.0 * E20[-1, 0] / (3 * n), T1, rtol=1e-12)


def test_thermostat_floor_at_zero(O):
    """beta = sqrt(max(0, 1 + gamma (T0/T - 1))): gamma > T/(T - T0) stops the atoms."""
    x, v = _free_gas(27, 7)
    p = O.o64_params(rc_lj=2.5)
    _, v1, _, E = O.vv_run64(x, v, None, None, p, 0.001, 12, thermostat_every=10,
                             temperature=0.0, gamma=2.0)
    assert np.all(v1 == 0.0) and E[-1, 0] == 0.0


def test_thermostat_period_zero_is_plain_vv(O):
    x, q, L = _liquid(343, 8)
    v = np.random.default_rng(8).normal(size=x.shape)
    p = O.o64_params(rc_lj=2.5, rc_coul=2.5)
    a = O.vv_run64(x, v, q, None, p, 0.001, 30)
    b = O.vv_run64(x, v, q, None, p, 0.001, 30, thermostat_every=0, temperature=5.0, gamma=1.0)
    for u, w in zip(a, b):
        assert np.array_equal(u, w)
