"""The host restatement of the tree (tigergrav_amd/csrc/tg_tree.cpp), no GPU needed.

Checks:
  * the particle order and node ranges equal an independent Python restatement of refine /
    part_vect_sort / box_id_to_range (src/tree.cpp:122-201, src/part_vect.cpp:895-933,
    src/range.cpp:134-169) -- the partition order is what fixes every later summation order;
  * leaves hold <= parts_per_node particles, children partition their parent, every particle lies
    in its leaf's box;
  * the lists of every leaf cover every particle exactly once, direct and Ewald (tests/
    list_invariants.py), and theta -> 0 leaves only particle-particle work;
  * evaluated exactly (no expansion), the lists with the reference's per-node periodic images
    reproduce the FP64 periodic sum -- the walk's image bookkeeping is consistent;
  * where the compiled reference exists, the oracle's whole solve over these lists is bit-identical
    to the same solve done by the reference's own L2L / L2P / kernels;
  * the solve's force error against the FP64 direct sum is the expansion's, not more.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from list_invariants import check_partition


def make_solver(x, ewald, theta=0.7, ppn=64, h=None):
    from tigergrav_amd.ics import soft_len
    from tigergrav_amd.solver import Solver

    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=h or soft_len(n), m=1.0 / n, theta=theta, parts_per_node=ppn, device=-1)
    s.set_particles(x)
    return s


# ---- independent restatement of the tree build (small N) ------------------------------------
def box_id_to_range(bid):
    x = [0.0, 0.0, 0.0]
    dx = [1.0, 1.0, 1.0]
    while bid != 1:
        x = [x[2], x[0], x[1]]
        dx = [dx[2], dx[0], dx[1]]
        x[0] = 0.5 * x[0] + 0.5 if bid & 1 else 0.5 * x[0]
        dx[0] *= 0.5
        bid >>= 1
    return x, [x[d] + dx[d] for d in range(3)]


def py_build(pos, ppn):
    pos = [list(p) for p in pos]
    order = list(range(len(pos)))
    p2d = lambda v: v * 2.0 ** -32 + 0.5
    ranges = []

    def refine(b, e, bid):
        ranges.append((b, e))
        if e - b <= ppn:
            return
        mn, mx = box_id_to_range(bid)
        span, dim = 0.0, 0
        for d in range(3):
            if mx[d] - mn[d] > span:
                span, dim = mx[d] - mn[d], d
        mid = (mx[dim] + mn[dim]) * 0.5
        lo, hi = b, e
        while lo < hi:
            if p2d(pos[lo][dim]) >= mid:
                while lo != hi:
                    hi -= 1
                    if p2d(pos[hi][dim]) < mid:
                        pos[lo], pos[hi] = pos[hi], pos[lo]
                        order[lo], order[hi] = order[hi], order[lo]
                        break
            lo += 1
        refine(b, hi, bid << 1)
        refine(hi, e, (bid << 1) + 1)

    refine(0, len(pos), 1)
    return np.array(pos, np.int32), np.array(order), sorted(ranges)


@pytest.mark.parametrize("n,ppn,gen", [(1, 1, "uniform"), (7, 1, "uniform"), (100, 4, "uniform"), (2000, 16, "uniform"),
                                       (3000, 64, "uniform"),
                                       # nodes above 2^16 particles are split by all threads at once
                                       (140000, 64, "uniform"), (140000, 128, "plummer")])
def test_partition_matches_restatement(n, ppn, gen):
    from tigergrav_amd import ics

    x = ics.uniform(n, seed=n) if gen == "uniform" else ics.plummer_spheres(n, seed=n)
    s = make_solver(x, ewald=True, ppn=ppn)
    pos, order, ranges = py_build(x, ppn)
    np.testing.assert_array_equal(s.export("pos"), pos)
    np.testing.assert_array_equal(s.export("order"), order)
    got = sorted(map(tuple, s.export("node_range")))
    assert got == ranges


@pytest.mark.parametrize("gen", ["uniform", "plummer", "zeldovich"])
def test_tree_structure(gen):
    from tigergrav_amd import ics

    x = {"uniform": lambda: ics.uniform(5000, 4), "plummer": lambda: ics.plummer_spheres(6000, 5),
         "zeldovich": lambda: ics.zeldovich(16, 6)}[gen]()
    s = make_solver(x, ewald=gen != "plummer", ppn=32)
    pos = s.export("pos")
    order = s.export("order")
    np.testing.assert_array_equal(pos, x[order])  # a permutation of the input
    rng, child, parent = s.export("node_range"), s.export("node_child"), s.export("node_parent")
    bid = s.export("node_boxid")
    for n in range(len(rng)):
        b, e = rng[n]
        if child[n][0] < 0:
            assert e - b <= 32
            if e > b:
                lo, hi = box_id_to_range(int(bid[n][0]) | (int(bid[n][1]) << 64))
                xd = pos[b:e] * 2.0 ** -32 + 0.5
                assert np.all(xd >= np.array(lo) - 1e-12) and np.all(xd <= np.array(hi) + 1e-12)
        else:
            c0, c1 = child[n]
            assert rng[c0][0] == b and rng[c0][1] == rng[c1][0] and rng[c1][1] == e
            assert parent[c0] == n and parent[c1] == n
    assert parent[0] == -1


@pytest.mark.parametrize("ewald,gen,theta", [(True, "uniform", 0.7), (False, "uniform", 0.7), (False, "plummer", 0.5),
                                             (True, "zeldovich", 0.7), (True, "uniform", 1.0)])
def test_lists_partition_all_particles(ewald, gen, theta):
    from tigergrav_amd import ics

    x = {"uniform": lambda: ics.uniform(4096, 7), "plummer": lambda: ics.plummer_spheres(6000, 8),
         "zeldovich": lambda: ics.zeldovich(16, 9)}[gen]()
    s = make_solver(x, ewald=ewald, theta=theta)
    T = s.tree_arrays()
    bad = check_partition(T, kinds=("d", "e") if ewald else ("d",))
    assert not bad, bad[:5]
    info = s.info()["count"]
    if not ewald:
        assert info["PP_ewald"] == info["CC_ewald"] == info["PC_ewald"] == info["CP_ewald"] == 0


def test_coincident_particles_beyond_leaf_size_stop_at_the_depth_limit():
    """More coincident particles than a leaf holds cannot be separated: the build stops splitting
    at level 60 and keeps them in one leaf, and every sink's lists still cover all particles once."""
    from tigergrav_amd import ics

    x = np.concatenate([np.repeat(ics.uniform(1, 5), 100, axis=0), ics.uniform(3000, 6)])
    s = make_solver(x, ewald=True, ppn=32)
    T = s.tree_arrays()
    r, lv = T["node_range"], T["node_level"]
    big = [lf for lf in T["leaves"] if r[lf][1] - r[lf][0] > 32]
    assert len(big) == 1 and r[big[0]][1] - r[big[0]][0] == 100 and lv[big[0]] == 60 and lv.max() == 60
    assert not check_partition(T, kinds=("d", "e"))


def test_tiny_theta_is_all_pp():
    from tigergrav_amd import ics

    s = make_solver(ics.uniform(1000, 3), ewald=True, theta=1e-3)
    c = s.info()["count"]
    assert c["PP_direct"] == 1000 * 1000 and c["PP_ewald"] == 1000 * 1000
    assert c["PC_direct"] == c["CC_direct"] == c["CP_direct"] == c["PC_ewald"] == c["CC_ewald"] == c["CP_ewald"] == 0


def test_periodic_images_are_consistent():
    """Summing every list item exactly (no expansions) with the node-centre minimum images the walk
    implies gives the FP64 periodic sum."""
    from tigergrav_amd import ics

    N = 1024
    x = ics.uniform(N, 11)
    h = ics.soft_len(N)
    s = make_solver(x, ewald=True, theta=0.7, ppn=32, h=h)
    T = s.tree_arrays()
    m = 1.0 / N
    P = T["pos"].astype(np.float64) * 2.0 ** -32 + 0.5
    NX = T["node_x"].astype(np.float64) * 2.0 ** -32 + 0.5
    rng = T["node_range"]
    w = lambda d: d - np.round(d)

    def grav(y):
        r = np.linalg.norm(y, axis=-1)
        ok = r > 0
        return -(y[ok] / r[ok, None] ** 3).sum(axis=0)

    def corr(y):
        return np.sum([ob.ewald_corr64(v)[1] for v in y], axis=0) if len(y) else np.zeros(3)

    leaf_of = np.empty(N, int)
    for lf in T["leaves"]:
        leaf_of[rng[lf][0]:rng[lf][1]] = lf
    sample = [3, 400, 777]
    _, g64 = ob.direct64(1, h, T["pos"][sample], T["pos"], m)
    for k, i in enumerate(sample):
        lf, xi, g = leaf_of[i], P[i], np.zeros(3)
        chain, n = [], lf
        while n >= 0:
            chain.append(n)
            n = T["node_parent"][n]
        for a in chain:
            Xa = NX[a]
            for kind, F in (("d", grav), ("e", corr)):
                for c in T["cc_" + kind][T["cc_" + kind + "_ptr"][a]:T["cc_" + kind + "_ptr"][a + 1]]:
                    ys, C = P[rng[c][0]:rng[c][1]], NX[c]
                    g += m * F(w(Xa - C) + (xi - Xa) - (ys - C))
                for b, e in T["cp_" + kind][T["cp_" + kind + "_ptr"][a]:T["cp_" + kind + "_ptr"][a + 1]].reshape(-1, 2):
                    g += m * F(w(Xa - P[b:e]) + (xi - Xa))
        for kind, F in (("d", grav), ("e", corr)):
            for c in T["pc_" + kind][T["pc_" + kind + "_ptr"][lf]:T["pc_" + kind + "_ptr"][lf + 1]]:
                ys, C = P[rng[c][0]:rng[c][1]], NX[c]
                g += m * F(w(xi - C) - (ys - C))
            for b, e in T["pp_" + kind][T["pp_" + kind + "_ptr"][lf]:T["pp_" + kind + "_ptr"][lf + 1]].reshape(-1, 2):
                g += m * F(w(xi - P[b:e]))
        # the particle's own softened self pair is zero force; the spline only matters below h
        assert np.linalg.norm(g - g64[k]) < 1e-9 * np.linalg.norm(g64[k]) + 1e-6, (g, g64[k])


@pytest.mark.skipif(not ob.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("ewald", [1, 0])
def test_oracle_solve_is_the_reference_solve(ewald):
    """Same lists, the reference's own code vs the oracle: bit-identical forces and flops."""
    import subprocess
    import sys

    code = f"""
import sys; sys.path.insert(0, {repr(ob.HERE)}); sys.path.insert(0, {repr(ob.os.path.dirname(ob.HERE))})
import numpy as np, oracle_bind as ob
from tigergrav_amd import ics
from tigergrav_amd.solver import Solver
N = 4096; x = ics.uniform(N, 1); h = ics.soft_len(N)
s = Solver(ewald=bool({ewald}), soft_len=h, m=1.0/N, theta=0.9, parts_per_node=64, device=-1); s.set_particles(x)
T = s.tree_arrays()
fo, a = ob.oracle_solve({{"ewald": {ewald}, "soft_len": h, "m": 1.0/N}}, T)
R = ob.ref_lib(); R.tgref_set_options({ewald}, h, 1.0, N, 0.9)
fr, b = ob.ref_solve(R, T)
assert a == b, (a, b)
assert np.array_equal(fo, fr), np.max(np.abs(fo - fr))
print("ok")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_solve_accuracy_open():
    """Open boundaries, theta = 0.7: rms relative force error of the oracle solve over the tree's
    lists against the FP64 direct sum."""
    from tigergrav_amd import ics

    N = 4096
    x = ics.uniform(N, 1)
    h = ics.soft_len(N)
    s = make_solver(x, ewald=False, theta=0.7, h=h)
    T = s.tree_arrays()
    f, _ = ob.oracle_solve({"ewald": 0, "soft_len": h, "m": 1.0 / N}, T)
    idx = np.arange(0, N, 16)
    _, g64 = ob.direct64(0, h, T["pos"][idx], T["pos"], 1.0 / N)
    e = np.linalg.norm(f[idx, 1:] - g64, axis=1) / np.linalg.norm(g64, axis=1)
    assert math.sqrt(np.mean(e ** 2)) < 3e-3


def test_update_particles_keeps_tree_and_recomputes_moments():
    """Frozen-tree update (tg_solver_update_particles): the same positions give bit-identical
    moments; moved particles keep the lists and expansion centres, and every leaf's moments are the
    raw moments about its kept centre (independent numpy P2M, float tolerance), every interior
    node's the M2M of its children's (tigergrav/multipole.hpp:414-433, restated here)."""
    from tigergrav_amd import ics

    n = 4096
    x = ics.uniform(n, 71)
    s = make_solver(x, True, ppn=32)
    m0 = s.export("node_m").copy()
    xc = s.export("node_x").copy()
    lists = {k: s.export(k).copy() for k in ("pp_d", "cc_d", "pc_d")}
    s.update_particles(x)
    assert np.array_equal(s.export("node_m"), m0)
    rng = np.random.default_rng(72)
    x2 = (x.astype(np.int64) + rng.integers(-2 ** 20, 2 ** 20, size=x.shape)).astype(np.int32)
    x2[:8, 0] = np.int32(2 ** 31 - 1)  # at the box edge ...
    x2[8:16, 0] = np.int32(-2 ** 31)   # ... and across it
    mass = rng.uniform(0.5, 1.5, n).astype(np.float32) / n
    s.update_particles(x2, mass)
    for k, v in lists.items():
        assert np.array_equal(s.export(k), v)
    assert np.array_equal(s.export("node_x"), xc)
    M = s.export("node_m").astype(np.float64)
    posi = s.export("pos").astype(np.int64)
    order = s.export("order")
    mt = mass[order].astype(np.float64)
    nr, ch = s.export("node_range"), s.export("node_child")
    X = xc.astype(np.float64) * 2.0 ** -32
    i2 = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    i3 = [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2), (0, 2, 2), (1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2)]
    for nd in range(len(nr)):
        b, e = nr[nd]
        if e == b:
            continue
        if ch[nd, 0] < 0:
            # periodic: offsets from the kept centre wrap around the box (some particles crossed its edge)
            d = (((posi[b:e] - xc[nd].astype(np.int64) + 2 ** 31) % 2 ** 32) - 2 ** 31) * 2.0 ** -32
            want = [mt[b:e].sum()] + [np.sum(mt[b:e] * d[:, p] * d[:, q]) for p, q in i2] + \
                   [np.sum(mt[b:e] * d[:, p] * d[:, q] * d[:, r]) for p, q, r in i3]
        else:
            want = np.zeros(17)
            for c in ch[nd]:
                y = X[c] - X[nd]
                mc = M[c]
                m2 = {pq: mc[1 + i] for i, pq in enumerate(i2)}
                s2 = lambda p, q: m2[(min(p, q), max(p, q))]
                want[0] += mc[0]
                for i, (p, q) in enumerate(i2):
                    want[1 + i] += mc[1 + i] + mc[0] * y[p] * y[q]
                for i, (p, q, r) in enumerate(i3):
                    want[7 + i] += mc[7 + i] + mc[0] * y[p] * y[q] * y[r] + s2(p, q) * y[r] + s2(q, r) * y[p] + s2(r, p) * y[q]
        want = np.asarray(want)
        for lo, hi in ((0, 1), (1, 7), (7, 17)):  # each order against its own magnitude
            scale = np.abs(want[lo:hi]).max() + 1e-30
            assert np.max(np.abs(M[nd, lo:hi] - want[lo:hi])) <= 2e-5 * scale, (nd, lo, M[nd], want)


def test_stage_times_are_reported():
    """tg_solve_info splits the host build into stages; a host-only solver reports its tree,
    multipole, walk and work-array preparation times (the device stages stay zero)."""
    from tigergrav_amd import ics

    s = make_solver(ics.uniform(5000, 17), ewald=True, ppn=32)
    i = s.info()
    assert i["t_tree"] > 0 and i["t_multipoles"] > 0 and i["t_walk"] > 0 and i["t_upload"] > 0
    d = i["t_detail"]
    assert set(d) == {"walk_setup", "walk_steps", "walk_assembly", "census", "prepare", "transfer"}
    assert d["prepare"] > 0 and d["prepare"] <= i["t_upload"]
    assert d["walk_steps"] == 0 and d["transfer"] == 0
