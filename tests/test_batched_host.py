"""Host side of the batched form: replica generation and spatial packing."""
import numpy as np


def test_replica_sizes_are_log_uniform_in_range(pkg):
    s = pkg.batched.replica_sizes(512, seed=11)
    assert len(s) == 512 and s.min() >= 256 and s.max() <= 16384
    # log-uniform: the median sits near the geometric mean of the range
    assert 1500 < np.median(s) < 2700
    assert 1.6e6 < s.sum() < 2.4e6


def test_pack_separates_replicas_by_more_than_the_gap(pkg):
    sizes = [256, 300, 517, 1024, 2000, 4096, 257]
    reps = pkg.batched.make_replicas(sizes, seed=3)
    gap = 3.8
    P = pkg.batched.pack(reps, gap, leaf=16)
    assert P["n"] % 16 == 0 and P["n"] - P["npad"] == sum(sizes)
    boxes = P["boxes"]
    # every atom inside its replica's box
    for r, R in enumerate(reps):
        a, b = P["start"][r], P["start"][r + 1]
        x = P["x"][a:b]
        assert np.all(x >= boxes[r, :3]) and np.all(x <= boxes[r, 3:])
        assert np.all(P["rep"][a:b] == r)
        assert np.allclose(x - P["origin"][r], R["x"])
    # boxes pairwise separated by at least `gap` along some axis
    for r in range(len(reps)):
        for t in range(r + 1, len(reps)):
            sep = np.maximum(boxes[t, :3] - boxes[r, 3:], boxes[r, :3] - boxes[t, 3:])
            assert sep.max() >= gap - 1e-9, (r, t, sep)
    # padding atoms: replica -1, far from every box and from each other
    pad = P["x"][P["n"] - P["npad"]:]
    assert np.all(P["rep"][P["n"] - P["npad"]:] == -1)
    for p in pad:
        d = np.maximum(np.maximum(boxes[:, :3] - p, p - boxes[:, 3:]), 0.0)
        assert np.linalg.norm(d, axis=1).min() >= gap
    if len(pad) > 1:
        assert np.min(np.linalg.norm(pad[1:] - pad[:-1], axis=1)) >= gap
    assert np.all(P["x"] >= 0.0) and np.all(P["x"] <= P["L"])
