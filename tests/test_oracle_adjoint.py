"""The FP64 adjoint oracle (oracle/adjoint64.c) against central finite
differences of its own forward model, and the RDF loss against its closed
form on a two-atom system.

The oracle is the checker of the GPU adjoint; these tests pin it: a dropped
term, a wrong sign or a wrong index in the reverse sweep changes the
gradient by far more than the finite-difference error (O(h^2) with h = 1e-5
in FP64 on a 60-step trajectory).
"""
import numpy as np
import pytest


def _system(O, n=48, seed=3, steps=60):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    x = g * 1.12 + rng.uniform(-0.08, 0.08, size=(n, 3))
    v = rng.normal(0, 1.0, size=(n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    L = side * 1.12
    p = O.Adj64Params(eps=1.0, sigma=1.0, rc=2.5, k_coul=1.0, rc_coul=2.5, dt=0.002,
                      rdf_rmax=2.5, rdf_h=0.15, volume=L ** 3, nbins=20, nsteps=steps)
    return x, v, q, p


def _target(O, x, v, q, p):
    pt = O.Adj64Params(*[getattr(p, f) for f, _ in p._fields_])
    pt.eps, pt.sigma = 1.1, 0.95
    _, g = O.adj64_run(x, v, q, pt)
    return g


def test_adjoint_oracle_matches_finite_differences(O):
    x, v, q, p = _system(O)
    t = _target(O, x, v, q, p)
    L0, g, dx, dv, dq, dpar = O.adj64_run(x, v, q, p, target=t, grads=True)
    assert L0 > 0.0
    h = 1e-5

    def loss(x_=x, v_=v, q_=q, eps=p.eps, sigma=p.sigma):
        pp = O.Adj64Params(*[getattr(p, f) for f, _ in p._fields_])
        pp.eps, pp.sigma = eps, sigma
        return O.adj64_run(x_, v_, q_, pp, target=t)[0]

    fd_eps = (loss(eps=p.eps + h) - loss(eps=p.eps - h)) / (2 * h)
    fd_sig = (loss(sigma=p.sigma + h) - loss(sigma=p.sigma - h)) / (2 * h)
    assert abs(dpar[0] - fd_eps) < 1e-6 * max(1.0, abs(fd_eps))
    assert abs(dpar[1] - fd_sig) < 1e-6 * max(1.0, abs(fd_sig))
    for i in (0, 7, 31):
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (loss(q_=qp) - loss(q_=qm)) / (2 * h)
        assert abs(dq[i] - fd) < 1e-6 * max(1.0, abs(fd)), (i, dq[i], fd)
    for (i, c) in ((0, 0), (5, 1), (40, 2)):
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        fd = (loss(x_=xp) - loss(x_=xm)) / (2 * h)
        assert abs(dx[i, c] - fd) < 1e-6 * max(1.0, abs(fd)), (i, c, dx[i, c], fd)
        vp, vm = v.copy(), v.copy()
        vp[i, c] += h
        vm[i, c] -= h
        fd = (loss(v_=vp) - loss(v_=vm)) / (2 * h)
        assert abs(dv[i, c] - fd) < 1e-6 * max(1.0, abs(fd)), (i, c, dv[i, c], fd)


def test_rdf_of_two_atoms_is_the_closed_form(O):
    # one pair at distance r: g_b = V / (1 * 4 pi r_b^2 dr) K(r - r_b)
    r = 1.234
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    v = np.zeros((2, 3))
    p = O.Adj64Params(eps=1.0, sigma=1.0, rc=2.5, k_coul=0.0, rc_coul=0.0, dt=1e-9,
                      rdf_rmax=2.5, rdf_h=0.2, volume=10.0, nbins=25, nsteps=1)
    _, g = O.adj64_run(x, v, None, p)
    dr = 0.1
    rb = (np.arange(25) + 0.5) * dr
    u = (r - rb) / 0.2
    K = np.where(np.abs(u) < 1, (1 - u * u) ** 3, 0.0)
    # the one step of dt = 1e-9 moves the atoms by ~1e-16
    assert np.allclose(g, 10.0 / (4 * np.pi * rb * rb * dr) * K, rtol=1e-9, atol=1e-12)


def _cluster(n, seed, L=4.0):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    x = g * (L / side) + rng.uniform(-0.1, 0.1, size=(n, 3))
    q = rng.permutation(np.r_[np.ones(n // 2), -np.ones(n - n // 2)])
    return x, q


def test_shifted_force_law_is_minus_the_energy_gradient(O):
    """The forward model's forces are -grad of U_sf(r) = U(r) - U(rc) - (r -
    rc) U'(rc) (energy from the potentials' definitions, central differences
    with h = 1e-6), and they vanish continuously at the cutoffs."""
    x, q = _cluster(27, 5)
    p = O.Adj64Params(eps=1.3, sigma=0.9, rc=2.5, k_coul=0.7, rc_coul=2.2, dt=0.001, rdf_rmax=2.0,
                      rdf_h=0.1, volume=64.0, nbins=10, nsteps=1)
    F, U = O.adj64_forces(x, q, p)
    h = 1e-6
    for (i, c) in ((0, 0), (4, 1), (13, 2), (26, 0)):
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        fd = -(O.adj64_forces(xp, q, p)[1] - O.adj64_forces(xm, q, p)[1]) / (2 * h)
        assert abs(F[i, c] - fd) < 1e-6 * max(1.0, abs(fd)), (i, c, F[i, c], fd)
    # a pair just inside and just beyond each cutoff: force and energy -> 0
    for r in (2.5 - 1e-7, 2.5 + 1e-7):
        x2 = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        F2, U2 = O.adj64_forces(x2, np.array([1.0, -1.0]), p)
        assert np.abs(F2).max() < 1e-6 and abs(U2) < 1e-12


@pytest.mark.parametrize("rc", [3.0, 10.0, 100.0])
def test_shifted_force_differs_from_textbook_by_the_cutoff_slope(O, rc):
    """Against the textbook forces of oracle64 (pinned by closed forms): for
    one pair, F_sf = F_textbook + U'(rc) r^ with U'_LJ(rc) = 4 eps (6 s^6 /
    rc^7 - 12 s^12 / rc^13) and U'_C(rc) = -k q1 q2 / rc^2, so the two agree
    to O(1/rc^2) in the large-cutoff limit."""
    eps, sig, k = 1.3, 0.9, 0.7
    q = np.array([1.0, -1.0])
    pa = O.Adj64Params(eps=eps, sigma=sig, rc=rc, k_coul=k, rc_coul=rc, dt=0.001, rdf_rmax=2.0,
                       rdf_h=0.1, volume=64.0, nbins=10, nsteps=1)
    pt = O.o64_params(eps=eps, sigma=sig, rc_lj=rc, k_coul=k, rc_coul=rc)
    for r in (0.95, 1.3, 2.2):
        x = np.array([[0.2, -0.1, 0.3], [0.2 + r * 0.6, -0.1 + r * 0.8, 0.3]])
        Fa, _ = O.adj64_forces(x, q, pa)
        Ft, _, _, _ = O.forces64(x, q, pt)
        rhat = (x[0] - x[1]) / r
        slope = 4 * eps * (6 * sig ** 6 / rc ** 7 - 12 * sig ** 12 / rc ** 13) - k * q[0] * q[1] / rc ** 2
        assert np.allclose(Fa[0], Ft[0] + slope * rhat, rtol=1e-11, atol=1e-13)
        assert np.allclose(Fa[1], -Fa[0], rtol=1e-12, atol=1e-14)
        assert np.abs(Fa[0] - Ft[0]).max() <= 1.01 * abs(slope)
