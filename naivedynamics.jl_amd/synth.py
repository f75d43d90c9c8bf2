"""Seeded synthetic inputs of the benchmark workloads (no data files).

Neighbour-search inputs are FP32 positions uniform in [0, L)^3 with L chosen
for reduced density 0.8.  Dynamics inputs cannot be plain uniform: at
density 0.8 uniform placement puts atoms at r << sigma and the LJ force
overflows on the first step.  Their initial-condition policy is a jittered
simple-cubic lattice: n_side = ceil(cbrt(N)) sites per edge at spacing
a = L / n_side, N sites drawn without replacement, each displaced uniformly in
[-0.1, 0.1)^3 (minimum separation a - 0.2 = 0.877 sigma at density 0.8).
Velocities are Gaussian with variance kT/m per component and zero mean
momentum; charges are +1/-1 in equal numbers on a random permutation.
All draws use numpy's PCG64 generator with the given seed.
"""
import numpy as np


def box_side(n, density=0.8):
    return float((n / density) ** (1.0 / 3.0))


def uniform_positions(n, L, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * L).astype(np.float32)


def lattice_positions(n, L, seed, jitter=0.1):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
    while side ** 3 < n:
        side += 1
    a = L / side
    sites = rng.choice(side ** 3, size=n, replace=False) if side ** 3 > n else np.arange(n)
    iz, rem = np.divmod(sites, side * side)
    iy, ix = np.divmod(rem, side)
    grid = np.stack([ix, iy, iz], axis=1).astype(np.float64)
    x = (grid + 0.5) * a + rng.uniform(-jitter, jitter, size=(n, 3))
    return x


def gaussian_velocities(n, kT, seed, mass=1.0):
    rng = np.random.default_rng(seed + 7919)
    v = rng.normal(0.0, np.sqrt(kT / mass), size=(n, 3))
    v -= v.mean(axis=0)
    return v


def balanced_charges(n, seed):
    rng = np.random.default_rng(seed + 104729)
    q = np.ones(n, np.float64)
    q[: n // 2] = -1.0
    return rng.permutation(q)


def plain_small(seed=1):
    """4,096 atoms, density 0.8, LJ only (eps = sigma = 1, rc = 2.5), dt = 0.001,
    100 steps, list rebuilt every step."""
    n = 4096
    L = box_side(n)
    return dict(workload="lj-4096", n=n, L=L, x=lattice_positions(n, L, seed),
                v=gaussian_velocities(n, 1.0, seed), q=None, rc=2.5, rc_coul=0.0, dt=0.001,
                steps=100, rebuild_every=1, skin=0.0)


def plain_scale(seed=2, n=262144):
    """262,144 atoms, density 0.8, LJ rc = 2.5 plus cutoff Coulomb (q = +-1,
    rc = 2.5), dt = 0.001, 1,000 steps, list rebuilt every 10 steps."""
    L = box_side(n)
    return dict(workload="ljc-262144", n=n, L=L, x=lattice_positions(n, L, seed),
                v=gaussian_velocities(n, 1.0, seed), q=balanced_charges(n, seed), rc=2.5,
                rc_coul=2.5, dt=0.001, steps=1000, rebuild_every=10, skin=0.3)


def blob_positions(n, L, seed, nblobs=64, std=2.0, spacing=0.95, jitter=0.05):
    """Clustered positions for the precision workload: nblobs centres uniform in
    [0, L)^3; the atoms take n sites of a simple-cubic lattice of the given
    spacing (jittered by +-jitter), chosen without replacement by Gumbel-top-k
    on the log density of the Gaussian mixture (one isotropic Gaussian of std
    `std` per centre).  Each blob therefore fills outwards from its centre at
    lattice density 1/spacing^3 with a soft edge, leaving voids between blobs,
    and no pair starts closer than spacing - 2 jitter."""
    rng = np.random.default_rng(seed)
    centres = rng.random((nblobs, 3)) * L
    side = int(np.floor(L / spacing))
    if side ** 3 < n:
        raise ValueError("blob_positions: lattice too coarse for n sites")
    a = L / side
    g = (np.arange(side) + 0.5) * a
    logw = np.empty(side ** 3)
    inv = -0.5 / (std * std)
    chunk = 1 << 18
    for s0 in range(0, side ** 3, chunk):
        idx = np.arange(s0, min(s0 + chunk, side ** 3))
        iz, rem = np.divmod(idx, side * side)
        iy, ix = np.divmod(rem, side)
        p = np.stack([g[ix], g[iy], g[iz]], axis=1)
        d2 = ((p[:, None, :] - centres[None, :, :]) ** 2).sum(axis=2) * inv
        m = d2.max(axis=1)
        logw[s0:s0 + len(idx)] = m + np.log(np.exp(d2 - m[:, None]).sum(axis=1))
    key = logw - np.log(-np.log(rng.random(side ** 3)))   # Gumbel-top-k
    sites = np.sort(np.argpartition(-key, n - 1)[:n])
    iz, rem = np.divmod(sites, side * side)
    iy, ix = np.divmod(rem, side)
    x = np.stack([g[ix], g[iy], g[iz]], axis=1) + rng.uniform(-jitter, jitter, size=(n, 3))
    return rng.permutation(x)


def precision_blobs(seed=3, n=1048576):
    """1,048,576 atoms in 64 blobs (box side for mean density 0.8, L = 109.4),
    LJ rc = 2.5 plus cutoff Coulomb (q = +-1, rc = 2.5), dt = 0.001, 500 steps,
    list rebuilt every 10 steps with skin 0.4.  The blobs start at lattice
    spacing 0.95 (compressed) and heat up as they relax: the mean speed goes
    from 1.6 to 3.4 over 500 steps and the fastest atoms reach |v| ~ 17, so an
    atom can move 0.17 in 10 steps; skin 0.4 (half of it 0.2) keeps the list
    valid over the whole run, where 0.3 would need forced rebuilds."""
    L = box_side(n)
    return dict(workload="ljc-blobs-1048576", n=n, L=L, x=blob_positions(n, L, seed),
                v=gaussian_velocities(n, 1.0, seed), q=balanced_charges(n, seed), rc=2.5,
                rc_coul=2.5, dt=0.001, steps=500, rebuild_every=10, skin=0.4)


__all__ = ["box_side", "uniform_positions", "lattice_positions", "gaussian_velocities",
           "balanced_charges", "plain_small", "plain_scale", "blob_positions", "precision_blobs"]
