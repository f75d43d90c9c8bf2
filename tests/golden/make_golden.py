"""Writes the golden fixtures in tests/golden/*.json.

An independent restatement of the reference's neighbour-search outputs in
pure Python + numpy, written from the definitions rather than from the
oracle's C code:
  * key:   q_d = floor(fl32(x_d / fl32(1/1023))), code = OR_d (q_d & M_d)
           (BVHTraverse.jl:184, 257-288);
  * order: numpy's stable argsort on int32 (Julia sortperm is stable);
  * tree:  built TOP-DOWN as the Cartesian tree of the delta sequence
           (BVHTraverse.jl:676-696): the range [l, r] splits after the right-most
           maximum of delta(l..r-1).  The reference builds bottom-up (Apetrei,
           :704-846); both give the same topology, so this checks the oracle's
           bottom-up CAS build against a different algorithm.  Node numbers,
           ropes and boxes follow from the topology: an internal node
           covering [l, r] is number nL + (delta(r) < delta(l-1) ? r : l); its
           rope is 0 if r = nL, else leaf r+1 if delta(r) < delta(r+1), else
           internal node nL+r+1; boxes are leaf +-r/2 boxes merged with Julia
           min/max; the root box is overwritten with [0,1]^3 (:856-858);
  * pairs: by definition of leafneighbor_traverse (:1236-1323): for query leaf
           q = 1..nL, same-leaf pairs (a < b), then every later leaf m > q whose
           box strictly overlaps q's box, tile in row-major order; accept
           s < fl32(r*r) (scalar) or fl32(sqrt(s)) < r (SIMD).  That is also the
           emission order of a single-threaded run.
The 8-point fixture is position8 of test/BVHTraverse.jl:165-169.

point_query_golden.json restates TreeData + neighbor_traverse
(BVHTraverse.jl:500-543, 1057-1092, 1125-1234) and the root-started form of
gpuneighbor_traverse! (ext/NaiveKA.jl:352-452) WITHOUT a tree walk.  Leaf
boxes are the atoms' +-r boxes merged with Julia min/max (cluster_primitives,
:341-374); an ancestor's box contains every leaf box below it, and a rope walk
from a leaf visits every later leaf whose ancestors it enters, so:
  * own-leaf form: for sorted q = 0..N-2 and sorted s = q+1..N-1 in order,
    keep (q, s) iff q lies strictly inside the box of s's leaf and
    s_qs < fl32(r*r);
  * root form: for sorted q = 0..N-1 and every sorted s in order, keep (q, s)
    iff perm[q] < perm[s], q strictly inside s's leaf box and s_qs < fl32(r*r);
pairs carry original indices and fl32(sqrt(s)).  store is the CAS-slot array
of a one-thread build: slot k (the split after leaf k) holds the first leaf of
the left child of the node that splits there.  quantized_positions!
(:290-314) is restated with Python's stable sort by (x, x is +0) per axis,
cascaded x then y then z.
Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F32 = np.float32
BW = F32(1.0 / 1023.0)
MAGIC = [0x09249249, 0x12492492, 0x24924924]
I32MAX, I32MIN = 2 ** 31 - 1, -2 ** 31


def wrap32(v):
    v &= 0xFFFFFFFF
    return v - 2 ** 32 if v >= 2 ** 31 else v


def masked_codes(x):
    x = np.asarray(x, F32)
    q = np.floor(x / BW).astype(np.int64)  # float32 division, then floor
    code = np.zeros(len(x), np.int64)
    for d in range(3):
        code |= q[:, d] & MAGIC[d]
    return code.astype(np.int32)


def jl_min(a, b):
    a, b = F32(a), F32(b)
    if a < b:
        return a
    if b < a:
        return b
    return a if np.signbit(a) else b


def jl_max(a, b):
    a, b = F32(a), F32(b)
    if a > b:
        return a
    if b > a:
        return b
    return b if np.signbit(a) else a


def build_topdown(xs, cs, apl, r, half=True, override_root=True, store=None):
    """Nodes in the reference numbering (1-based), as list of dicts.  Leaf
    boxes are +-r/2 (half) or +-r; store (dict) receives split -> first leaf
    of the left child."""
    N = len(xs)
    nL = N // apl
    h = F32(r) / F32(2) if half else F32(r)

    def delta(i):
        if i >= nL or i < 1:
            return I32MAX
        x = int(cs[(i - 1) * apl]) ^ int(cs[i * apl])
        if x == 0:
            return wrap32(x + (I32MIN + (i ^ (i + 1))) - 1)
        return wrap32(x - 1)

    nodes = {}
    for l in range(1, nL + 1):
        pts = xs[(l - 1) * apl: l * apl]
        mn = [F32(pts[0][d] - h) for d in range(3)]
        mx = [F32(pts[0][d] + h) for d in range(3)]
        for p in pts:
            for d in range(3):
                mn[d] = jl_min(mn[d], F32(p[d] - h))
                mx[d] = jl_max(mx[d], F32(p[d] + h))
        nodes[l] = {"min": mn, "max": mx, "left": 0, "skip": 0}

    def rope(rr):
        if rr == nL:
            return 0
        return rr + 1 if delta(rr) < delta(rr + 1) else rr + 1 + nL

    def make(l, r):
        if l == r:
            nodes[l]["skip"] = rope(l)
            return l
        ds = [delta(k) for k in range(l, r)]
        m = max(ds)
        k = l + max(i for i, v in enumerate(ds) if v == m)
        left = make(l, k)
        right = make(k + 1, r)
        if store is not None:
            store[k] = l
        idx = nL + (r if delta(r) < delta(l - 1) else l)
        mn = [jl_min(nodes[right]["min"][d], nodes[left]["min"][d]) for d in range(3)]
        mx = [jl_max(nodes[right]["max"][d], nodes[left]["max"][d]) for d in range(3)]
        nodes[idx] = {"min": mn, "max": mx, "left": left, "skip": rope(r)}
        return idx

    root = make(1, nL)
    assert root == nL + 1
    if override_root:
        nodes[root]["min"] = [F32(0), F32(0), F32(0)]
        nodes[root]["max"] = [F32(1), F32(1), F32(1)]
    return [nodes[k] for k in range(1, 2 * nL)]


def sqdist(a, b):
    dx, dy, dz = F32(a[0] - b[0]), F32(a[1] - b[1]), F32(a[2] - b[2])
    s = F32(F32(dx * dx) + F32(dy * dy))
    return F32(s + F32(dz * dz))


def pairs_by_definition(nodes, xs, perm, apl, r, simd):
    nL = len(xs) // apl
    r = F32(r)
    r2 = F32(r * r)
    out = []

    def acc(s):
        if simd:
            d = F32(np.sqrt(s))
            return (d < r), d
        return (s < r2), F32(np.sqrt(s))

    def overlap(A, B):
        return all((A["min"][d] < B["max"][d]) and (A["max"][d] > B["min"][d]) for d in range(3))

    for q in range(1, nL + 1):
        lo = (q - 1) * apl
        for a in range(apl - 1):
            for b in range(a + 1, apl):
                ok, d = acc(sqdist(xs[lo + a], xs[lo + b]))
                if ok:
                    out.append((int(perm[lo + a]), int(perm[lo + b]), float(d)))
        for m in range(q + 1, nL + 1):
            if not overlap(nodes[q - 1], nodes[m - 1]):
                continue
            mlo = (m - 1) * apl
            for a in range(apl):
                for b in range(apl):
                    ok, d = acc(sqdist(xs[lo + a], xs[mlo + b]))
                    if ok:
                        out.append((int(perm[lo + a]), int(perm[mlo + b]), float(d)))
    return out


def brute_force(x, r):
    out = []
    r = F32(r)
    for i in range(len(x) - 1):
        for j in range(i + 1, len(x)):
            d = F32(np.sqrt(sqdist(x[i], x[j])))
            if d < r:
                out.append((i, j, float(d)))
    return out


def case(name, x, apl, r):
    x = np.asarray(x, F32)
    codes = masked_codes(x)
    perm = np.argsort(codes, kind="stable").astype(np.int32)
    xs = x[perm]
    cs = codes[perm]
    nodes = build_topdown(xs, cs, apl, r)
    enc = lambda v: [int(np.asarray(F32(t)).view(np.uint32)) for t in v]  # noqa: E731
    return {
        "name": name, "apl": apl, "r": float(F32(r)),
        "position_bits": [enc(p) for p in x],
        "codes": [int(c) for c in codes],
        "perm": [int(p) for p in perm],
        "nodes": [{"min_bits": enc(n["min"]), "max_bits": enc(n["max"]), "left": n["left"],
                   "skip": n["skip"]} for n in nodes],
        "pairs_scalar": [[i, j, int(np.asarray(F32(d)).view(np.uint32))]
                         for i, j, d in pairs_by_definition(nodes, xs, perm, apl, r, False)],
        "pairs_simd": [[i, j, int(np.asarray(F32(d)).view(np.uint32))]
                       for i, j, d in pairs_by_definition(nodes, xs, perm, apl, r, True)],
        "bruteforce": [[i, j, int(np.asarray(F32(d)).view(np.uint32))]
                       for i, j, d in brute_force(x, r)],
    }


def point_query_by_definition(xs, perm, apl, r, root_form):
    """Point-query pairs from the +-r leaf boxes alone (no tree)."""
    N = len(xs)
    nL = N // apl
    r = F32(r)
    r2 = F32(r * r)
    boxes = []
    for l in range(nL):
        pts = xs[l * apl:(l + 1) * apl]
        mn = [F32(pts[0][d] - r) for d in range(3)]
        mx = [F32(pts[0][d] + r) for d in range(3)]
        for pt in pts:
            for d in range(3):
                mn[d] = jl_min(mn[d], F32(pt[d] - r))
                mx[d] = jl_max(mx[d], F32(pt[d] + r))
        boxes.append((mn, mx))
    out = []
    for q in range(N if root_form else N - 1):
        p = xs[q]
        for s in (range(N) if root_form else range(q + 1, N)):
            if root_form and not perm[q] < perm[s]:
                continue
            mn, mx = boxes[s // apl]
            if not all(mn[d] < p[d] < mx[d] for d in range(3)):
                continue
            d2 = sqdist(p, xs[s])
            if d2 < r2:
                out.append((int(perm[q]), int(perm[s]), float(F32(np.sqrt(d2)))))
    return out


def ranks_by_definition(x):
    """quantized_positions!: cascaded stable sorts, isless order (-0 < +0)."""
    N = len(x)
    order = list(range(N))
    q = [[0, 0, 0] for _ in range(N)]
    for d in range(3):
        order = sorted(order, key=lambda i: (float(x[i][d]), not np.signbit(x[i][d])))
        for k, i in enumerate(order):
            q[i][d] = k + 1
    return q


def pq_case(name, x, apl, r):
    x = np.asarray(x, F32)
    codes = masked_codes(x)
    perm = np.argsort(codes, kind="stable").astype(np.int32)
    xs = x[perm]
    cs = codes[perm]
    store = {}
    nodes = build_topdown(xs, cs, apl, r, half=False, override_root=False, store=store)
    enc = lambda v: [int(np.asarray(F32(t)).view(np.uint32)) for t in v]  # noqa: E731
    nL = len(x) // apl
    pairs = lambda root: [[i, j, int(np.asarray(F32(d)).view(np.uint32))]  # noqa: E731
                          for i, j, d in point_query_by_definition(xs, perm, apl, r, root)]
    return {
        "name": name, "apl": apl, "r": float(F32(r)),
        "position_bits": [enc(p) for p in x],
        "perm": [int(p) for p in perm],
        "nodes_root_kept": [{"min_bits": enc(n["min"]), "max_bits": enc(n["max"]),
                             "left": n["left"], "skip": n["skip"]} for n in nodes],
        "store_serial": [int(store[k]) for k in range(1, nL)],
        "pairs_own_leaf": pairs(False),
        "pairs_root": pairs(True),
        "quantized": ranks_by_definition(x),
    }


def margin_pair(r=0.1, seed=7):
    """Two atoms a = fl(c - r) and c on one axis with fl((a - c)^2) < fl(r*r):
    within the radius, yet a sits on the boundary of c's +-r box."""
    rng = np.random.default_rng(seed)
    r = F32(r)
    for _ in range(100000):
        c = F32(rng.uniform(0.3, 0.7))
        a = F32(c - r)
        if sqdist([a, 0.5, 0.5], [c, 0.5, 0.5]) < F32(r * r):
            return [[float(a), 0.5, 0.5], [float(c), 0.5, 0.5]]
    raise RuntimeError("no margin pair found")


def main():
    position8 = [[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.346, 0.98, 0.12], [0.01, 0.76, 0.99],
                 [0.1111, 0.4, 0.31], [0.234, 0.29, 0.2], [0.11346, 0.918, 0.1276],
                 [0.061, 0.76, 0.989]]
    cases = [case("position8_apl1_r10", position8, 1, 10.0),
             case("position8_apl2_r0.5", position8, 2, 0.5),
             case("position8_apl4_r0.7", position8, 4, 0.7)]
    rng = np.random.default_rng(20250101)
    cases.append(case("unit64_apl4_r0.3", rng.random((64, 3)), 4, 0.3))
    cases.append(case("unit200_apl5_r0.15", rng.random((200, 3)), 5, 0.15))
    L = (256 / 0.8) ** (1 / 3)
    cases.append(case("rho08_256_apl4_r2.8", rng.random((256, 3)) * L, 4, 2.8))
    with open(os.path.join(HERE, "neighbor_golden.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print("wrote", len(cases), "cases")

    rng = np.random.default_rng(20261203)
    pq = [pq_case("position8_apl1_r10", position8, 1, 10.0),
          pq_case("position8_apl2_r0.5", position8, 2, 0.5),
          pq_case("unit128_apl4_r0.2", rng.random((128, 3)), 4, 0.2),
          pq_case("unit200_apl5_r0.15", rng.random((200, 3)), 5, 0.15)]
    # ties and signed zeros for the rank sorts; two margin pairs (boundary of
    # the +-r box, inside the radius) among uniform points
    ties = rng.choice(np.array([-0.0, 0.0, 0.25, -0.5, 0.75], F32), size=(60, 3))
    ties[:, 0] = np.abs(ties[:, 0]) * 0.5 + rng.random(60).astype(F32) * 0.25
    pq.append(pq_case("ties60_apl2_r0.3", ties, 2, 0.3))
    m = margin_pair() + margin_pair(seed=11)
    pts = np.concatenate([np.asarray(m, F32), rng.random((60, 3)).astype(F32)])
    pq.append(pq_case("margin64_apl1_r0.1", pts, 1, 0.1))
    # beyond the unit cube: the own-leaf walk never meets the overwritten root
    pq.append(pq_case("box3_96_apl4_r0.6", rng.random((96, 3)) * 3.0, 4, 0.6))
    with open(os.path.join(HERE, "point_query_golden.json"), "w") as f:
        json.dump(pq, f, separators=(",", ":"))
    print("wrote", len(pq), "point-query cases")


if __name__ == "__main__":
    main()
