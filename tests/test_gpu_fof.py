"""Friends-of-friends groups on the device (tg_solver_find_groups, include/tg_fof.h) against
  * the oracle's restatement of the reference's relabelling (tgo_fof) on small sets, bit for bit
    (group ids are tree-order indices, as the reference's);
  * at 2M particles (8 Plummer spheres and a background), an independent partition: scipy's k-d tree for candidate pairs, the
    reference's integer test on them, connected components, each labelled by its smallest
    tree-order index;
and the group catalog (tg_solver_group_catalog, src/groups.cpp) against numpy statistics of the same
members: counts exact, centre, mean velocity, dispersion, largest and median radius to rounding."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from tigergrav_amd import ics

pytestmark = pytest.mark.gpu
NONE = 0xFFFFFFFFFFFFFF


class Group(C.Structure):
    _fields_ = [("id", C.c_uint64), ("N", C.c_int32), ("rmax", C.c_float), ("rc", C.c_float), ("x", C.c_double * 3),
                ("v", C.c_float * 3), ("dv", C.c_float * 3)]


def _groups(s, link_len):
    from tigergrav_amd._lib import lib

    L = lib()
    L.tg_solver_find_groups.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_int64)]
    g = np.zeros(s.n, np.uint64)
    ng = C.c_int64()
    rc = L.tg_solver_find_groups(s._h, link_len, g.ctypes.data, C.byref(ng))
    assert rc == 0, L.tg_solver_last_error()
    return g, ng.value


def _catalog(s, v, a, min_n):
    from tigergrav_amd._lib import lib

    L = lib()
    L.tg_solver_group_catalog.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_void_p, C.c_int64,
                                          C.POINTER(C.c_int64)]
    n = C.c_int64()
    assert L.tg_solver_group_catalog(s._h, v.ctypes.data, a, min_n, None, 0, C.byref(n)) == 0
    out = (Group * max(n.value, 1))()
    assert L.tg_solver_group_catalog(s._h, v.ctypes.data, a, min_n, out, n.value, C.byref(n)) == 0
    return list(out)[: n.value]


def _solver(x, ewald, leaf=32):
    from tigergrav_amd.solver import Solver

    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.7, parts_per_node=leaf, device=0)
    s.set_particles(x)
    return s


@pytest.mark.parametrize("ewald", [True, False])
def test_fof_matches_oracle(ewald):
    x = ics.plummer_spheres(3000, 31) if not ewald else np.concatenate([ics.plummer_spheres(2000, 32), ics.uniform(1000, 33)])
    n = x.shape[0]
    print(f"fof vs oracle: N={n} ewald={ewald}", flush=True)
    s = _solver(x, ewald)
    g, ng = _groups(s, 0.2)
    T = s.tree_arrays()
    want_tree = ob.fof(T["pos"], ob.fof_link2(n, 0.2))
    got_tree = g[T["order"]]
    assert np.array_equal(got_tree, want_tree)
    assert ng == int(np.sum(want_tree == np.arange(n, dtype=np.uint64)))
    print(f"  {ng} groups, {int(np.sum(g != NONE))} linked particles", flush=True)


def _reference_partition(pos_tree, link_len):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree

    n = pos_tree.shape[0]
    L2 = ob.fof_link2(n, link_len)
    xd = pos_tree.astype(np.float64) * 2.0 ** -32 + 0.5
    kd = cKDTree(xd)  # not periodic: a pair across the box edge is not a friend
    pairs = kd.query_pairs(np.sqrt(L2) / 2 ** 32 * (1 + 1e-6) + 1e-12, output_type="ndarray")
    d = (pos_tree[pairs[:, 0]].astype(np.int64) - pos_tree[pairs[:, 1]].astype(np.int64)).astype(np.uint64)
    d2 = np.sum(d * d, axis=1, dtype=np.uint64)
    keep = (d2 < np.uint64(L2)) & (d2 != 0)
    pairs = pairs[keep]
    A = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    nc, lab = connected_components(A, directed=False)
    first = np.full(nc, n, np.int64)
    np.minimum.at(first, lab, np.arange(n))
    linked = np.zeros(n, bool)
    linked[pairs[:, 0]] = True
    linked[pairs[:, 1]] = True
    return np.where(linked, first[lab].astype(np.uint64), np.uint64(NONE))


def test_fof_2m_and_catalog():
    x = ics.plummer_spheres(2 ** 21, seed=2)
    n = x.shape[0]
    print(f"fof plummer2M: N={n}", flush=True)
    s = _solver(x, False)
    import time

    t0 = time.perf_counter()
    g, ng = _groups(s, 0.2)
    print(f"  find_groups {time.perf_counter() - t0:.3f} s, {ng} groups", flush=True)
    T = s.tree_arrays()
    order = T["order"]
    want = _reference_partition(T["pos"], 0.2)
    assert np.array_equal(g[order], want)
    # catalog of groups with >= 10 members against numpy on the same members
    rng = np.random.default_rng(9)
    v = rng.standard_normal((n, 3)).astype(np.float32)
    a = 0.5
    cat = _catalog(s, v, a, 10)
    ids, counts = np.unique(want[want != NONE], return_counts=True)
    big = ids[counts >= 10]
    assert [c.id for c in cat] == [int(i) for i in big]
    gt = want
    vt = v[order].astype(np.float64) / a
    xt = T["pos"].astype(np.float64) * 2.0 ** -32 + 0.5
    for c in cat[:: max(1, len(cat) // 50)]:
        m = gt == np.uint64(c.id)
        assert c.N == int(m.sum())
        # centre: the mean of the members unwrapped about one of them (compact groups)
        xm = xt[m]
        u = xm[0] + (np.mod(xm - xm[0] + 0.5, 1.0) - 0.5)
        cen = np.mod(u.mean(0), 1.0)
        dc = np.abs(np.array(c.x) - cen)
        assert np.all(np.minimum(dc, 1 - dc) < 1e-9), (c.id, c.x[:], cen)
        vm = vt[m].mean(0)
        np.testing.assert_allclose(np.array(c.v[:]), vm, rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(np.array(c.dv[:]), vt[m].std(0), rtol=1e-3, atol=1e-5)
        dx = np.abs(xt[m] - np.array(c.x))
        dx = np.minimum(dx, 1 - dx)
        r = a * np.sqrt(np.sum(dx * dx, axis=1))
        assert abs(c.rmax - r.max()) <= 1e-6 * max(r.max(), 1e-30) + 1e-12
        assert abs(c.rc - np.sort(r)[len(r) // 2]) <= 1e-6 * max(r.max(), 1e-30) + 1e-12
    print(f"  catalog: {len(cat)} groups of >= 10", flush=True)
