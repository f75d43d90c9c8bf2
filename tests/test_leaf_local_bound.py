"""CPU: the two-phase walk's leaf-local pair test and its rounding bound.

The walk tests an atom p against a staged leaf atom q as
    t = fma(a.x, -2q'.x, fma(a.y, -2q'.y, fma(a.z, -2q'.z, |q'|^2))) + cp,
    a = p - c, q' = q - c (c the leaf's centre), cp = |a|^2 - rl2,
all in FP32, and takes the exact test's result (|p - q|^2 < rl2 computed
from the coordinate differences) whenever |t| < E = 48 u (cp + 2 rl2 + R^2),
u = 2^-24, R^2 the leaf's largest |q'|^2.  The list stays the exact test's
only if every t with |t| >= E has the exact test's sign.  Here that is
checked on FP32 emulations (FMA = one rounding of the exact float64 value)
of boundary pairs, loose leaves and far-off coordinates; E is also checked
to exceed the measured deviation by a margin (the largest seen is ~0.07 E),
and a bound smaller by 64x is shown to fail (so the check can fail)."""
import numpy as np

F = np.float32


def fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F)


def exact_s(p, q):
    # the walk's exact test: dy*dy, then + dx*dx, then + dz*dz (FMAs)
    d = (p - q).astype(F)
    s = fma(d[:, 2], d[:, 2], fma(d[:, 0], d[:, 0], (d[:, 1] * d[:, 1]).astype(F)))
    return s


def local_test(p, q, c, R2, rl2):
    a = (p - c).astype(F)
    qq = (q - c).astype(F)
    q2 = fma(qq[:, 2], qq[:, 2], fma(qq[:, 1], qq[:, 1], (qq[:, 0] * qq[:, 0]).astype(F)))
    m2 = (F(-2.0) * qq).astype(F)
    cp = fma(a[:, 2], a[:, 2], fma(a[:, 1], a[:, 1], fma(a[:, 0], a[:, 0], np.full(len(a), -rl2, F))))
    E = (F(48.0 * 2.0 ** -24) * ((cp + F(2.0) * rl2).astype(F) + R2).astype(F)).astype(F)
    t = (fma(a[:, 0], m2[:, 0], fma(a[:, 1], m2[:, 1], fma(a[:, 2], m2[:, 2], q2))) + cp).astype(F)
    return t, E, q2


def _cases(rng, m, scale, half, rl):
    """m leaves of centre up to `scale`, half-extent `half`: an atom q in the
    leaf and p at a distance within a few ulp of rl from q (or anywhere within
    rl + 2 half of the centre)."""
    c = rng.uniform(0.0, scale, (m, 3)).astype(F)
    q = (c + rng.uniform(-half, half, (m, 3))).astype(F)
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    near = rng.random(m) < 0.8
    dist = np.where(near, rl * (1.0 + rng.uniform(-4e-7, 4e-7, m)), rng.uniform(0.0, rl + 2 * half, m))
    p = (q + u * dist[:, None]).astype(F)
    # the leaf's R^2: this atom's |q'|^2 and up to the corner of its box
    R2 = np.full(m, F(3.0 * half * half), F)
    return p, q, c, R2


def test_local_test_sign_agrees_outside_the_band():
    rng = np.random.default_rng(5)
    rl = 2.8
    rl2 = F(rl * rl)
    worst = 0.0
    for scale, half in ((70.0, 2.0), (70.0, 6.0), (1000.0, 2.0), (250.0, 12.0)):
        p, q, c, R2 = _cases(rng, 200000, scale, half, rl)
        t, E, q2 = local_test(p, q, c, R2, rl2)
        assert np.all(q2 <= R2)
        s = exact_s(p, q)
        dev = np.abs(t.astype(np.float64) - (s.astype(np.float64) - np.float64(rl2)))
        worst = max(worst, float(np.max(dev / E)))
        out = np.abs(t) >= E
        assert np.array_equal((t < 0)[out], (s < rl2)[out])
        # the pairs within a few ulp of the radius (80 %) fall in the band,
        # where the exact test decides; the others are decided by t
        assert np.mean(~out) < 0.9
    # the bound holds with room to spare
    assert worst < 0.25, worst


def test_a_much_smaller_bound_would_misorder_pairs():
    rng = np.random.default_rng(6)
    rl = 2.8
    rl2 = F(rl * rl)
    p, q, c, R2 = _cases(rng, 200000, 70.0, 2.0, rl)
    t, E, _ = local_test(p, q, c, R2, rl2)
    s = exact_s(p, q)
    out = np.abs(t) >= E / 64
    assert not np.array_equal((t < 0)[out], (s < rl2)[out])
