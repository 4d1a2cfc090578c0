"""GPU batched force solve against the oracle's whole-solve restatement on the same lists.

The oracle solve (tgo_solve) is the reference's solve sequence -- L2L, L2P and the eight
kernels, node by node in kick_fmm's order -- with the bit-faithful float kernels; on these lists it
is bit-identical to the same solve done by the compiled reference (tests/test_tree.py).
Tolerance: rms relative error of accelerations and of potentials <= 1e-5.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

CASES = {
    # name: (generator, ewald, theta, parts_per_node)
    "uniform_periodic": (lambda ics: ics.uniform(2048, 21), True, 0.7, 32),
    "uniform_open": (lambda ics: ics.uniform(4096, 22), False, 0.7, 64),
    "plummer_open": (lambda ics: ics.plummer_spheres(8192, 23), False, 0.7, 64),
    "zeldovich_periodic": (lambda ics: ics.zeldovich(12, 24), True, 0.7, 32),
    "uniform_periodic_leaf128": (lambda ics: ics.uniform(2048, 25), True, 1.0, 128),
}


EDGE_CASES = {
    # name: (generator, ewald, theta, parts_per_node): one particle (zero force by symmetry), two,
    # a single leaf, a ragged count with leaves of 16, and 40 coincident particles (r = 0 pairs)
    "n1_periodic": (lambda ics: ics.uniform(1, 91), True, 0.7, 32),
    "n2_open": (lambda ics: ics.uniform(2, 92), False, 0.7, 32),
    "n63_periodic_one_leaf": (lambda ics: ics.uniform(63, 93), True, 0.7, 64),
    "ragged_1001_periodic": (lambda ics: ics.uniform(1001, 94), True, 0.5, 16),
    "coincident_cluster": (lambda ics: np.concatenate([np.repeat(ics.uniform(1, 95), 40, axis=0), ics.uniform(2000, 96)]),
                           True, 0.7, 64),
    # 100 coincident particles with leaves of 32: one 100-particle leaf at the depth limit
    "coincident_beyond_leaf": (lambda ics: np.concatenate([np.repeat(ics.uniform(1, 97), 100, axis=0), ics.uniform(3000, 98)]),
                               True, 0.7, 32),
}


def rms_rel(a, b):
    return math.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2))


@pytest.mark.parametrize("case", list(CASES))
def test_solve_parity(case):
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    gen, ewald, theta, ppn = CASES[case]
    x = gen(ics)
    N = x.shape[0]
    h, m = ics.soft_len(N), 1.0 / N
    print(f"case {case}: N={N} ewald={ewald} theta={theta} leaf={ppn}", flush=True)
    s = Solver(ewald=ewald, soft_len=h, m=m, theta=theta, parts_per_node=ppn, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    T = s.tree_arrays()
    ref_tree, _ = ob.oracle_solve({"ewald": int(ewald), "soft_len": h, "m": m}, T)
    ref = np.empty_like(ref_tree)
    ref[T["order"]] = ref_tree
    eg = rms_rel(f[:, 1:], ref[:, 1:])
    ep = rms_rel(f[:, 0], ref[:, 0])
    print(f"  rms rel g {eg:.2e}  phi {ep:.2e}", flush=True)
    assert eg < 1e-5 and ep < 1e-5
    # a second solve on the resident data is identical
    s.solve()
    s.synchronize()
    np.testing.assert_array_equal(s.forces(), f)
    ms = s.kernel_ms()
    assert all(v >= 0 for v in ms.values())


@pytest.mark.parametrize("chunk", [64, 256, 500, 4096])
def test_solve_parity_long_chunks(chunk):
    """CC/CP partial slots of more than 64 list entries: one wave per slot whose lanes each add
    several entries in order before the store.  Periodic Zel'dovich set at theta 0.3 (node lists of
    hundreds of entries) against the oracle solve on the same lists, at the parity tolerance."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.zeldovich(24, 27)
    N = x.shape[0]
    h, m = ics.soft_len(N), 1.0 / N
    print(f"long chunks {chunk}: N={N} theta=0.3 leaf=16", flush=True)
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.3, parts_per_node=16, device=0)
    s.tune("node_chunk", chunk)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    T = s.tree_arrays()
    cc = np.diff(T["cc_d_ptr"])
    assert cc.max() > min(chunk, 400)  # slots longer than a wave occur
    ref_tree, _ = ob.oracle_solve({"ewald": 1, "soft_len": h, "m": m}, T)
    ref = np.empty_like(ref_tree)
    ref[T["order"]] = ref_tree
    eg = rms_rel(f[:, 1:], ref[:, 1:])
    dphi = math.sqrt(np.mean((f[:, 0] - ref[:, 0]) ** 2))
    print(f"  rms rel g {eg:.2e}  rms |dphi| {dphi:.2e} (total mass 1); longest CC list {cc.max()}", flush=True)
    # the periodic potential's rounding is bounded in absolute terms, as for the full-size test
    assert eg < 1e-5 and dphi < 2e-6
    s.solve_graph()
    s.synchronize()
    np.testing.assert_array_equal(s.forces(), f)


@pytest.mark.parametrize("case", list(EDGE_CASES))
def test_solve_edge_cases(case):
    """Small, ragged and coincident particle sets against the oracle solve on the same lists:
    every component within 2e-5 of the case's largest (a zero force stays zero)."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    gen, ewald, theta, ppn = EDGE_CASES[case]
    x = np.ascontiguousarray(gen(ics))
    N = x.shape[0]
    h, m = ics.soft_len(N), 1.0 / N
    print(f"edge case {case}: N={N} ewald={ewald} theta={theta} leaf={ppn}", flush=True)
    s = Solver(ewald=ewald, soft_len=h, m=m, theta=theta, parts_per_node=ppn, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    T = s.tree_arrays()
    ref_tree, _ = ob.oracle_solve({"ewald": int(ewald), "soft_len": h, "m": m}, T)
    ref = np.empty_like(ref_tree)
    ref[T["order"]] = ref_tree
    gs = max(np.abs(ref[:, 1:]).max(), 1e-30)
    ps = np.abs(ref[:, 0]).max()
    assert np.abs(f[:, 1:] - ref[:, 1:]).max() <= 2e-5 * gs or np.abs(f[:, 1:]).max() == 0.0 == np.abs(ref[:, 1:]).max()
    assert np.abs(f[:, 0] - ref[:, 0]).max() <= 2e-5 * ps
    s.solve_graph()
    s.synchronize()
    np.testing.assert_array_equal(s.forces(), f)


def test_empty_particle_set():
    """No particles: the solve, its graph form and the force readback all succeed and return nothing;
    particles set afterwards on the same solver solve as usual."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    print("empty particle set", flush=True)
    s = Solver(ewald=True, soft_len=1e-3, m=1.0, theta=0.7, parts_per_node=32, device=0)
    s.set_particles(np.zeros((0, 3), np.int32))
    s.solve()
    s.solve_graph()
    s.synchronize()
    assert s.forces().shape == (0, 4)
    x = ics.uniform(1000, 99)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    assert f.shape == (1000, 4) and np.all(np.isfinite(f))


def test_solve_accuracy_vs_fp64():
    """Force error of the GPU solve against the FP64 periodic direct sum matches the expansion
    error of the reference algorithm (the oracle solve on the same lists)."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    N = 2048
    x = ics.uniform(N, 31)
    h, m = ics.soft_len(N), 1.0 / N
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.5, parts_per_node=32, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    idx = np.arange(0, N, 16)
    _, g64 = ob.direct64(1, h, x[idx], x, m)
    e = np.linalg.norm(f[idx, 1:] - g64, axis=1) / np.linalg.norm(g64, axis=1)
    T = s.tree_arrays()
    ref_tree, _ = ob.oracle_solve({"ewald": 1, "soft_len": h, "m": m}, T)
    ref = np.empty_like(ref_tree)
    ref[T["order"]] = ref_tree
    e_ref = np.linalg.norm(ref[idx, 1:] - g64, axis=1) / np.linalg.norm(g64, axis=1)
    rms, rms_ref = math.sqrt(np.mean(e ** 2)), math.sqrt(np.mean(e_ref ** 2))
    print(f"rms |da|/|a| gpu {rms:.3e} reference algorithm {rms_ref:.3e}")
    assert abs(rms - rms_ref) < 1e-4 + 1e-3 * rms_ref


def test_solve_parity_full_size():
    """The 2,097,152-particle periodic workload of the benchmark (128^3 Zel'dovich, theta 0.7, leaf
    64) against the oracle solve on the same lists.  Accelerations: rms relative <= 1e-5.  The
    periodic potential is a small difference of O(1) terms (the Ewald constant pi/4 per source
    mass, the self terms), so its single-precision rounding is bounded in absolute terms: rms
    |dphi| <= 2e-6 times the total mass (6.7e-7 measured)."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.zeldovich(128, seed=3)
    N = x.shape[0]
    h, m = ics.soft_len(N), 1.0 / N
    print(f"full size: N={N} periodic theta=0.7 leaf=64", flush=True)
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.7, parts_per_node=64, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    f = s.forces()
    T = s.tree_arrays()
    ref_tree, _ = ob.oracle_solve({"ewald": 1, "soft_len": h, "m": m}, T)
    ref = np.empty_like(ref_tree)
    ref[T["order"]] = ref_tree
    eg = rms_rel(f[:, 1:], ref[:, 1:])
    dphi = math.sqrt(np.mean((f[:, 0] - ref[:, 0]) ** 2))
    print(f"  rms rel g {eg:.2e}  rms |dphi| {dphi:.2e} (total mass 1)", flush=True)
    assert eg < 1e-5 and dphi < 2e-6


def test_graph_replay_matches_eager():
    """The solve replayed from a captured HIP graph gives the eager solve's forces bit for bit, on
    every replay, with eager solves and host synchronisation in between, and after the particles
    change (the graph is captured again)."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    for n, seed in ((4096, 51), (3000, 52)):
        x = ics.uniform(n, seed)
        print(f"graph replay: N={n} seed={seed}", flush=True)
        s = Solver(ewald=True, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.7, parts_per_node=32, device=0)
        s.set_particles(x)
        s.solve()
        s.synchronize()
        eager = s.forces()
        for rep in range(4):
            s.solve_graph()
            s.synchronize()
            np.testing.assert_array_equal(s.forces(), eager)
            s.solve()  # eager work between replays
            s.synchronize()
            np.testing.assert_array_equal(s.forces(), eager)
        # new particle set on the same solver: the graph is rebuilt for the new buffers
        x2 = ics.uniform(n // 2, seed + 100)
        s.set_particles(x2)
        s.solve()
        s.synchronize()
        eager2 = s.forces()
        for rep in range(3):
            s.solve_graph()
            s.synchronize()
            np.testing.assert_array_equal(s.forces(), eager2)


def test_overlap_knob_is_bit_identical():
    """Graph replays with the node kinds on a second stream ("overlap" 1, the default), with the
    Ewald node kinds on a third ("overlap" 2) and on one stream give the eager solve's forces bit
    for bit, with both local-pass forms."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.zeldovich(27, 71)
    n = x.shape[0]
    print(f"overlap knob: N={n}", flush=True)
    s = Solver(ewald=True, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.6, parts_per_node=32, device=0)
    s.set_particles(x)
    for sweep in (0, 1):
        s.tune("local_sweep", sweep)
        s.solve()
        s.synchronize()
        eager = s.forces()
        for ov in (1, 0, 2, 1, 2):
            s.tune("overlap", ov)
            for _ in range(2):
                s.solve_graph()
                s.synchronize()
                np.testing.assert_array_equal(s.forces(), eager)


def test_tuning_knobs():
    """The persistent queue-driven PP kernel evaluates each item exactly as the one-group-per-item
    kernel does (bit-identical forces, any residency); a different CC/CP chunk only regroups FP64
    partial sums (within the parity tolerance of the default); the level-by-level local pass
    differs from the path sums by FP64 rounding only."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    n = 8192
    x = ics.plummer_spheres(n, 61)
    h, m = ics.soft_len(n), 1.0 / n
    print(f"tuning knobs: N={n} periodic", flush=True)
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.7, parts_per_node=64, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    base = s.forces()
    for resident in (1, 7, 5000):
        s.tune("pp_persistent", 1)
        s.tune("resident", resident)
        s.solve()
        s.synchronize()
        np.testing.assert_array_equal(s.forces(), base)
        s.solve_graph()
        s.synchronize()
        np.testing.assert_array_equal(s.forces(), base)
    for chunk in (16, 200):
        s2 = Solver(ewald=True, soft_len=h, m=m, theta=0.7, parts_per_node=64, device=0)
        s2.tune("node_chunk", chunk)
        s2.set_particles(x)
        s2.solve()
        s2.synchronize()
        f2 = s2.forces()
        assert rms_rel(f2[:, 1:], base[:, 1:]) < 1e-6
    # the level-by-level local pass shifts the running sum node by node (the reference's order):
    # the same expansion up to FP64 rounding, and graph replays of it bit-identical to eager solves
    s.tune("pp_persistent", 0)
    s.tune("local_sweep", 1)
    s.solve()
    s.synchronize()
    f3 = s.forces()
    assert rms_rel(f3[:, 1:], base[:, 1:]) < 1e-12 and np.abs(f3[:, 0] - base[:, 0]).max() < 1e-12 * np.abs(base[:, 0]).max()
    for _ in range(3):
        s.solve_graph()
        s.synchronize()
        np.testing.assert_array_equal(s.forces(), f3)
    with pytest.raises(Exception):
        s2.tune("no_such_knob", 1)


def test_rebuild_is_bit_identical():
    """Two solvers built from the same particles give bit-identical forces: the host build runs
    in parallel (tree levels, multipoles, work arrays) and the lists come from the device walk,
    but every array they produce is fixed by the particles alone."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.plummer_spheres(200000, 71)
    n = x.shape[0]
    out = []
    for leaf, walk in ((32, 1), (32, 1), (32, 0)):
        s = Solver(ewald=True, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.6, parts_per_node=leaf, device=0)
        s.tune("gpu_walk", walk)
        s.set_particles(x)
        s.solve()
        s.synchronize()
        out.append(s.forces())
        s.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_autotune_picks_the_fastest_tried_configuration():
    """The autotuner times graph replays of every configuration it tries and returns the fastest;
    its choice applied to a fresh solver gives the same forces as the eager solve of that
    configuration."""
    from tigergrav_amd import ics
    from tigergrav_amd.autotune import apply, autotune
    from tigergrav_amd.solver import Solver

    x = ics.uniform(6000, 81)
    n = x.shape[0]
    h, m = ics.soft_len(n), 1.0 / n
    print(f"autotune: N={n}", flush=True)
    best, table = autotune(x, True, h, m, 0.7, leaves=(32, 64), chunks=(64,), solves=3)
    assert len(table) == 2 * 1 * 2 * 2 * 3  # leaves x chunks x PP forms x local passes x overlap
    assert all(r["ms"] > 0 for r in table) and best["ms"] == min(r["ms"] for r in table)
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.7, parts_per_node=best["leaf"], device=0)
    apply(s, best)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    eager = s.forces()
    s.solve_graph()
    s.synchronize()
    np.testing.assert_array_equal(s.forces(), eager)


def test_update_moves_pair_inside_h_across_far_range():
    """After update_particles, a particle moved to within h/4 of a particle of a leaf that its own
    leaf lists as PP but classed far at the build (leaf boxes more than 2h apart) is still summed
    with the softened pair: the frozen-tree solve equals the oracle solve (which softens every
    listed pair) on the same lists, for both leaves' particles."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    n = 4096
    x = ics.uniform(n, 41)
    h, m = ics.soft_len(n), 1.0 / n
    s = Solver(ewald=True, soft_len=h, m=m, theta=0.7, parts_per_node=32, device=0)
    s.set_particles(x)
    T = s.tree_arrays()
    pos = T["pos"] * 2.0 ** -32 + 0.5
    rng_, p, v = T["node_range"], s.export("pp_d_ptr"), s.export("pp_d").reshape(-1, 2)
    box = {}
    for lf in T["leaves"]:
        b, e = rng_[lf]
        if e > b:
            box[b] = (pos[b:e].min(0), pos[b:e].max(0), lf)
    pick = None
    for A in T["leaves"]:
        b, e = rng_[A]
        if e <= b:
            continue
        for sb, se in v[p[A] // 2:p[A + 1] // 2]:
            mnB, mxB, B = box[sb]
            mnA, mxA = pos[b:e].min(0), pos[b:e].max(0)
            gap = np.maximum(0.0, np.maximum(mnB - mxA, mnA - mxB))
            if np.sqrt(np.sum(gap ** 2)) > 2 * h:
                pick = (A, B, b, sb)
                break
        if pick:
            break
    assert pick is not None
    A, B, i_tree, j_tree = pick
    print(f"update: leaf {A} particle moved to h/4 of leaf {B}'s particle", flush=True)
    order = T["order"]
    x2 = x.copy()
    off = np.array([int(h / 4 * 2 ** 32), 0, 0], np.int64)
    x2[order[i_tree]] = (T["pos"][j_tree].astype(np.int64) + off).astype(np.int32)
    s.update_particles(x2)
    s.solve()
    s.synchronize()
    f = s.forces()
    T2 = s.tree_arrays()
    ref_tree, mask = ob.oracle_solve_leaves({"ewald": 1, "soft_len": h, "m": m}, T2, [A, B])
    f_tree = f[T2["order"]]
    eg = rms_rel(f_tree[mask, 1:], ref_tree[mask, 1:])
    ei = np.linalg.norm(f_tree[i_tree, 1:] - ref_tree[i_tree, 1:]) / np.linalg.norm(ref_tree[i_tree, 1:])
    print(f"  rms rel g over both leaves {eg:.2e}, moved particle {ei:.2e}", flush=True)
    assert eg < 1e-5 and ei < 1e-5
