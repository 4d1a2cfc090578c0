"""The friends-of-friends checker (tgo_fof, oracle/tg_oracle.c: part_vect_find_groups of
src/part_vect.cpp:84-120 repeated to convergence) pinned by hand-made sets and by an independent
partition: a numpy union-find over all pairs within the linking length.

The rules fixed by the reference: friends are pairs whose squared int32 wrap-around separation, in
uint64, is below L^2 and not zero; a linked particle ends with the smallest index of its connected
component; a particle with no friend keeps DEFAULT_GROUP (2^56 - 1).  A pair across the box edge is
not a friend (the reference's candidate ranges are not wrapped)."""
import numpy as np

import oracle_bind as ob
from kernel_cases import double_to_pos

NONE = 0xFFFFFFFFFFFFFF


def pos(xd):
    return np.asarray(double_to_pos(np.asarray(xd, np.float64)), np.int32).reshape(-1, 3)


def test_hand_sets():
    L = 0.01
    L2 = int(L * 2 ** 32) ** 2
    # a chain 0-1-2 with gaps below L, a pair 3-4, a loner 5, a coincident pair 6-7 (distance 0:
    # not friends), and 8, 9 within L only across the periodic edge (not friends)
    xd = [[0.5, 0.5, 0.5], [0.509, 0.5, 0.5], [0.518, 0.5, 0.5],
          [0.2, 0.2, 0.2], [0.2, 0.205, 0.2],
          [0.8, 0.1, 0.1],
          [0.3, 0.7, 0.3], [0.3, 0.7, 0.3],
          [0.999, 0.4, 0.4], [0.004, 0.4, 0.4]]
    g = ob.fof(pos(xd), L2)
    assert list(g) == [0, 0, 0, 3, 3, NONE, NONE, NONE, NONE, NONE]
    # order matters only for the ids: reversing the particles relabels each group by its new
    # smallest index
    g2 = ob.fof(pos(xd[::-1]), L2)
    assert list(g2) == [NONE, NONE, NONE, NONE, NONE, 5, 5, 7, 7, 7]


def test_partition_matches_union_find():
    rng = np.random.default_rng(3)
    n = 600
    centres = rng.uniform(0, 1, (12, 3))
    xd = np.mod(centres[rng.integers(0, 12, n)] + rng.normal(0, 0.01, (n, 3)), 1.0)
    xd[:40, 0] = rng.uniform(0.998, 1.0, 40)  # some near the edge, with partners across it
    xd[40:80, 0] = rng.uniform(0.0, 0.002, 40)
    x = pos(xd)
    L2 = ob.fof_link2(n, 0.2)
    g = ob.fof(x, L2)
    # independent: union-find over the same integer test
    d = (x[:, None, :].astype(np.int64) - x[None, :, :].astype(np.int64))
    near = np.all(np.abs(d) < 2 ** 31, axis=2)  # no wrap: a pair across the edge is not a friend
    du = np.where(near[..., None], d, 0).astype(np.uint64)  # two's complement, as vect<uint64_t>
    d2 = np.sum(du * du, axis=2, dtype=np.uint64)
    link = near & (d2 < np.uint64(L2)) & (d2 != 0)
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            a = parent[a]
        return a

    for i, j in zip(*np.nonzero(np.triu(link, 1))):
        a, b = find(i), find(j)
        if a != b:
            parent[max(a, b)] = min(a, b)
    want = np.array([find(i) if link[i].any() else NONE for i in range(n)], np.uint64)
    assert np.array_equal(g, want)
    assert len(set(want[want != NONE].tolist())) > 5
