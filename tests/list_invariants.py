"""Completeness of the interaction lists (helper for the tree tests).

For every sink leaf, the sources it sees through the direct lists -- CC/CP items of itself and all
its ancestors plus its own PC and PP items -- must cover every particle exactly once; the same
holds for the Ewald lists.  This is what makes the walk (src/tree.cpp:378-635) a partition of
the N^2 interactions.
"""
import numpy as np


def coverage(T, leaf, kind="d"):
    """Per-particle multiplicity of the sources seen by `leaf` through the `kind` lists."""
    N = T["pos"].shape[0]
    rng = T["node_range"]
    cnt = np.zeros(N + 1, np.int64)

    def add_range(b, e):
        cnt[b] += 1
        cnt[e] -= 1

    n = leaf
    chain = []
    while n >= 0:
        chain.append(n)
        n = T["node_parent"][n]
    for a in chain:
        for name, is_range in (("cc_" + kind, False), ("cp_" + kind, True)):
            p, v = T[name + "_ptr"], T[name]
            seg = v[p[a]:p[a + 1]]
            if is_range:
                for b, e in seg.reshape(-1, 2):
                    add_range(b, e)
            else:
                for c in seg:
                    add_range(*rng[c])
    for name, is_range in (("pc_" + kind, False), ("pp_" + kind, True)):
        p, v = T[name + "_ptr"], T[name]
        seg = v[p[leaf]:p[leaf + 1]]
        if is_range:
            for b, e in seg.reshape(-1, 2):
                add_range(b, e)
        else:
            for c in seg:
                add_range(*rng[c])
    return np.cumsum(cnt)[:N]


def check_partition(T, leaves=None, kinds=("d",)):
    bad = []
    leaves = T["leaves"] if leaves is None else leaves
    for lf in leaves:
        b, e = T["node_range"][lf]
        if b == e:
            continue
        for k in kinds:
            c = coverage(T, lf, k)
            if not np.all(c == 1):
                bad.append((int(lf), k, int(np.sum(c == 0)), int(np.sum(c > 1))))
    return bad
