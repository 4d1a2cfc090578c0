"""An FP64 check of interaction lists against the reference's opening criterion, independent of the
list builders (tg_tree.cpp's walk_rec, tg_walk.hip): it builds no lists, it takes exported lists
and node data and asserts, entry by entry, what the walk of src/tree.cpp:378-635 implies for every
entry it can emit.

The far test (src/tree.cpp:429,456,537,574), for a sink node a and a source node c:
    dX = pos_to_double(x_a) - pos_to_double(x_c)       (FP64, per component)
    direct:  dx = ewald_near_separation(dX) when periodic, |dX| otherwise   (src/gravity.cpp:10-18)
    Ewald:   dx = max(0.25, ewald_near_separation(dX))                      (src/gravity.cpp:20-23)
    far  <=> dx > (r_a + r_c) / theta  and  dx > (r_a + r_c) + h   with r_a + r_c a float sum
    Ewald list with dX == 0:  far <=> rmax_a < r_ewald, rmax_a the span of a's particle box
    (src/tree.cpp:406) and r_ewald = float(2 theta (1e-3 / 8)^(1/3)) (src/tree.cpp:403-404).

What the walk implies for each list entry (sink a, source c; p = parent(c); A a leaf):
  CC (a, c): far(a, c), and c's parent was opened at a's parent: not far(parent(a), p).
  CP (a, c): c is a leaf, far(a, c); it was carried down opened, so not far(parent(a), c).
  PC (A, c): far(A, c); c came from opening p in A's own drain: not far(A, p).
  PP (A, c): c is a leaf and not far(A, c); p was opened somewhere on A's root path.
Together with the coverage invariant (tests/list_invariants.py: every sink sees every particle
exactly once) this pins which entries exist.  It does not pin their order.
"""
import numpy as np

POS_INV = 2.0 ** -32


def _p2d(x):
    return np.asarray(x, np.float64) * POS_INV + 0.5


def _near_sep(dX):
    a = np.abs(dX)
    t = np.minimum(a, 1.0 - a)
    return np.sqrt((t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1]) + t[:, 2] * t[:, 2])


def _abs(dX):
    return np.sqrt((dX[:, 0] * dX[:, 0] + dX[:, 1] * dX[:, 1]) + dX[:, 2] * dX[:, 2])


class MacVerifier:
    def __init__(self, T, theta, h, ewald):
        """T: exported tree arrays (Solver.tree_arrays() plus the lists)."""
        self.T = T
        self.theta_inv = 1.0 / theta
        self.h = float(h)
        self.ewald = bool(ewald)
        self.r_ewald = np.float32(2.0 * theta * (1.0e-3 / 8.0) ** (1.0 / 3.0))
        self.X = _p2d(T["node_x"])
        self.r = np.asarray(T["node_r"], np.float32)
        self.parent = np.asarray(T["node_parent"], np.int64)
        self.child = np.asarray(T["node_child"], np.int64)
        self.rng = np.asarray(T["node_range"], np.int64)
        self.is_leaf = self.child[:, 0] < 0
        self.sinks = None
        # span of each node's particle box, from the particles themselves: leaves by a scan of
        # their particles, other nodes as the union of their non-empty children, deepest level first
        pos = _p2d(T["pos"])
        nn = self.rng.shape[0]
        ne = self.rng[:, 1] > self.rng[:, 0]
        mx = np.full((nn, 3), -np.inf)
        mn = np.full((nn, 3), np.inf)
        lv = np.nonzero(self.is_leaf & ne)[0]
        if len(lv):
            o = lv[np.argsort(self.rng[lv, 0], kind="stable")]
            b = self.rng[o, 0]
            assert np.all(self.rng[o[:-1], 1] == b[1:]), "leaves must tile the particles"
            mx[o] = np.stack([np.maximum.reduceat(pos[:, d], b) for d in range(3)], 1)
            mn[o] = np.stack([np.minimum.reduceat(pos[:, d], b) for d in range(3)], 1)
        level = np.asarray(T["node_level"], np.int64)
        for l in range(int(level.max()) if nn else 0, -1, -1):
            k = np.nonzero((level == l) & ~self.is_leaf)[0]
            for j in (0, 1):
                c = self.child[k, j]
                mx[k] = np.maximum(mx[k], mx[c])
                mn[k] = np.minimum(mn[k], mn[c])
        s = np.where(ne[:, None], mx - mn, 0.0)
        self.rmax = np.sqrt((s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2]).astype(np.float32)
        # leaf lookup by first particle (non-empty leaves have distinct first particles)
        lv = np.nonzero(self.is_leaf & ne)[0]
        self._leaf_b = self.rng[lv, 0]
        o = np.argsort(self._leaf_b, kind="stable")
        self._leaf_b, self._leaf_id = self._leaf_b[o], lv[o]

    # ---- the far test, vectorised over (a, c) pairs --------------------------------------------
    def far(self, a, c, kind):
        a = np.asarray(a, np.int64)
        c = np.asarray(c, np.int64)
        dX = self.X[a] - self.X[c]
        rsum = (self.r[a] + self.r[c]).astype(np.float32).astype(np.float64)
        if kind == "d":
            dx = _near_sep(dX) if self.ewald else _abs(dX)
            return (dx > rsum * self.theta_inv) & (dx > rsum + self.h)
        dx = np.maximum(0.25, _near_sep(dX))
        mac = (dx > rsum * self.theta_inv) & (dx > rsum + self.h)
        zero = ((dX[:, 0] * dX[:, 0] + dX[:, 1] * dX[:, 1]) + dX[:, 2] * dX[:, 2]) == 0.0
        return np.where(zero, self.rmax[a] < self.r_ewald, mac)

    def leaf_of_range(self, b, e):
        k = np.searchsorted(self._leaf_b, b)
        ok = (k < len(self._leaf_b))
        k = np.minimum(k, len(self._leaf_b) - 1)
        c = self._leaf_id[k]
        ok &= (self._leaf_b[k] == b) & (self.rng[c, 1] == e)
        return c, ok

    def _entries(self, name, ranges):
        p = np.asarray(self.T[name + "_ptr"], np.int64)
        v = np.asarray(self.T[name], np.int64)
        per = 2 if ranges else 1
        counts = np.diff(p) // per
        sink = np.repeat(np.arange(len(counts)), counts)
        if ranges:
            v = v.reshape(-1, 2)
        if self.sinks is not None:
            keep = self.sinks[sink]
            sink, v = sink[keep], v[keep]
        if ranges:
            c, ok = self.leaf_of_range(v[:, 0], v[:, 1])
            return sink, c, ok
        return sink, v, np.ones(len(v), bool)

    def _on_path_not_far(self, A, p, kind):
        """for each (A, p): p not far from some node on A's root path (A included)"""
        out = np.zeros(len(A), bool)
        a = np.asarray(A, np.int64).copy()
        live = np.ones(len(A), bool)
        while live.any():
            idx = np.nonzero(live)[0]
            out[idx] |= ~self.far(a[idx], p[idx], kind)
            a[idx] = self.parent[a[idx]]
            live[idx] = a[idx] >= 0
        return out

    def check(self, kinds=None, sink_leaves=None):
        """Counts of violated conditions per list (all zero for lists the reference's walk builds).
        With sink_leaves, only the entries of those leaves and of their ancestors are checked."""
        kinds = kinds or (("d", "e") if self.ewald else ("d",))
        self.sinks = None
        if sink_leaves is not None:
            self.sinks = np.zeros(len(self.parent), bool)
            a = np.asarray(sink_leaves, np.int64)
            while len(a):
                self.sinks[a] = True
                a = self.parent[a]
                a = np.unique(a[a >= 0])
        bad = {}
        for k in kinds:
            # CC: far, and the parent of c opened at the parent of a
            a, c, _ = self._entries("cc_" + k, False)
            f = self.far(a, c, k)
            pa, pc = self.parent[a], self.parent[c]
            has = (pa >= 0) & (pc >= 0)
            opened = np.ones(len(a), bool)
            opened[has] = ~self.far(pa[has], pc[has], k)
            bad["cc_" + k] = int(np.sum(~f | ~opened | ~has))
            # CP: a leaf, far, carried down opened (not far at a's parent)
            a, c, ok = self._entries("cp_" + k, True)
            f = self.far(a, c, k)
            pa = self.parent[a]
            carried = np.zeros(len(a), bool)
            carried[pa >= 0] = ~self.far(pa[pa >= 0], c[pa >= 0], k)
            bad["cp_" + k] = int(np.sum(~ok | ~f | ~carried))
            # PC: at a leaf, far, its parent opened in that leaf's drain
            A, c, _ = self._entries("pc_" + k, False)
            f = self.far(A, c, k)
            pc = self.parent[c]
            drained = np.zeros(len(A), bool)
            drained[pc >= 0] = ~self.far(A[pc >= 0], pc[pc >= 0], k)
            bad["pc_" + k] = int(np.sum(~self.is_leaf[A] | ~f | ~drained))
            # PP: at a leaf, a leaf source, not far; its parent opened on the sink's root path
            A, c, ok = self._entries("pp_" + k, True)
            f = self.far(A, c, k)
            pc = self.parent[c]
            root_src = pc < 0
            opened = np.ones(len(A), bool)
            nz = ~root_src
            opened[nz] = self._on_path_not_far(A[nz], pc[nz], k)
            bad["pp_" + k] = int(np.sum(~ok | ~self.is_leaf[A] | f | ~opened))
        return bad


def coverage_exact(T, leaves, kind):
    """Leaves (of `leaves`) whose `kind` lists do not tile [0, N) exactly once: the ranges seen by a
    leaf through its own PC/PP lists and its root path's CC/CP lists, sorted, must abut from 0 to N."""
    N = T["pos"].shape[0]
    rng = np.asarray(T["node_range"], np.int64)
    parent = np.asarray(T["node_parent"], np.int64)
    P = {n: (np.asarray(T[n + "_ptr"], np.int64), np.asarray(T[n], np.int64))
         for n in ("cc_" + kind, "cp_" + kind, "pc_" + kind, "pp_" + kind)}

    def seg(name, a, ranges):
        p, v = P[name]
        s = v[p[a]:p[a + 1]]
        return s.reshape(-1, 2) if ranges else rng[s]

    bad = []
    for lf in leaves:
        if rng[lf, 1] == rng[lf, 0]:
            continue
        parts = [seg("pc_" + kind, lf, False), seg("pp_" + kind, lf, True)]
        a = lf
        while a >= 0:
            parts += [seg("cc_" + kind, a, False), seg("cp_" + kind, a, True)]
            a = parent[a]
        r = np.concatenate(parts)
        r = r[r[:, 1] > r[:, 0]]
        r = r[np.argsort(r[:, 0], kind="stable")]
        ok = len(r) > 0 and r[0, 0] == 0 and r[-1, 1] == N and np.all(r[1:, 0] == r[:-1, 1])
        if not ok:
            bad.append(int(lf))
    return bad
