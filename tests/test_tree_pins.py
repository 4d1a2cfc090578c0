"""Pins of the tree and the interaction lists that do not come from this build's own builders.

1. Hand-computed trees (no GPU).  For a few particle sets of N <= 5, the particle order, the node
   ranges by box id, the centres of mass, radii and moments below were worked out by hand from the
   reference's rules, not by running any code:
     * split the longest box side, the first dimension on equal spans (src/tree.cpp:136-145), at
       the box midpoint (src/tree.cpp:155-157); box ranges from box ids (src/range.cpp:134-169);
     * the in-place partition of src/part_vect.cpp:905-926 (a value equal to the midpoint is high);
     * COM = N-weighted mean of the children's quantised COMs (src/tree.cpp:240-244); radius
       r = min(max_child(|dCOM| + r_child), farthest corner of the particle box) (src/tree.cpp:
       249-275); moments raw about the COM, no dipole (src/part_vect.cpp:567-635, M2M
       tigergrav/multipole.hpp:117-135);
     * the lists of the walk (src/tree.cpp:378-635), for an open 4-particle set at theta 0.7 and
       0.5, in the reference's order.
2. The FP64 MAC verifier (tests/mac_verify.py) and exact coverage on host-built lists of several
   workloads, among them the 32^3 uniform periodic set at theta 0.7, leaf 64; and checks that the
   verifier fails lists built with another theta or softening.
"""
import numpy as np
import pytest

from kernel_cases import double_to_pos
from mac_verify import MacVerifier, coverage_exact


def solver_for(xd, ewald=False, theta=0.7, ppn=1, h=1e-3, m=None):
    from tigergrav_amd.solver import Solver

    x = np.asarray(double_to_pos(np.asarray(xd, np.float64)), np.int32).reshape(-1, 3)
    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=h, m=m or 1.0 / n, theta=theta, parts_per_node=ppn, device=-1)
    s.set_particles(x)
    return s


def by_boxid(s):
    bid = s.export("node_boxid")
    ids = [int(lo) | (int(hi) << 64) for lo, hi in bid]
    return ids, {b: i for i, b in enumerate(ids)}


def ranges_by_boxid(s):
    ids, _ = by_boxid(s)
    return {b: tuple(int(v) for v in r) for b, r in zip(ids, s.export("node_range"))}


# ---- 1. hand-computed trees -------------------------------------------------------------------
H1 = [(0.9, 0.3, 0.3), (0.1, 0.7, 0.3), (0.6, 0.3, 0.7), (0.3, 0.2, 0.3), (0.2, 0.8, 0.6)]
H2 = [(0.5, 0.5, 0.5), (0.25, 0.5, 0.5), (0.75, 0.25, 0.5), (0.5, 0.75, 0.25)]
H3 = [(0.2, 0.6, 0.5), (0.3, 0.7, 0.5), (0.8, 0.5, 0.5)]


def test_hand_tree_h1():
    # root: equal spans -> x, mid 0.5; x = .9 .1 .6 .3 .2: lo 0 high, swap with 4; lo 2 high, swap
    # with 3; stop at 3 -> order 4 1 3 2 0, left [0,3), right [3,5) (a leaf of 2).
    # left box [0,.5]x[0,1]^2 -> y (first of the two spans of 1), mid .5; y = .8 .7 .2: swap 0 and
    # 2; lo 1 high, hi falls to 1 -> split at 1: [3] and [1 4].
    s = solver_for(H1, ppn=2)
    assert list(s.export("order")) == [3, 1, 4, 2, 0]
    assert ranges_by_boxid(s) == {1: (0, 5), 2: (0, 3), 3: (3, 5), 4: (0, 1), 5: (1, 3)}


def test_hand_tree_h2_ties_and_box_ranges():
    # root x, mid .5; x = .5 .25 .75 .5: p0 equals the midpoint (high), swapped with p1 -> split at 1.
    # right box [.5,1]x[0,1]^2 -> y: y(p0,p2,p3) = .5 .25 .75 -> swap p0, p2 -> [p2] [p0 p3].
    # box 7 = [.5,1]x[.5,1]x[0,1] -> z: z(p0,p3) = .5 .25 -> swap -> [p3] [p0].
    s = solver_for(H2, ppn=1)
    assert list(s.export("order")) == [1, 2, 3, 0]
    assert ranges_by_boxid(s) == {1: (0, 4), 2: (0, 1), 3: (1, 4), 6: (1, 2), 7: (2, 4), 14: (2, 3), 15: (3, 4)}
    from tigergrav_amd.solver import Solver  # noqa: F401  (box_id_to_range through the exported ids)


def test_hand_tree_h3_empty_children():
    # root x: .2 .3 .8 -> [p0 p1] [p2].  box 2 -> y, mid .5, y = .6 .7: both high, left empty.
    # box 5 = [0,.5]x[.5,1]x[0,1] -> z, z = .5 .5 (ties, high): left empty again.
    # box 11 = [0,.5]x[.5,1]x[.5,1] -> x (equal spans of .5), mid .25: x = .2 .3 -> [p0] [p1].
    s = solver_for(H3, ppn=1)
    assert list(s.export("order")) == [0, 1, 2]
    assert ranges_by_boxid(s) == {1: (0, 3), 2: (0, 2), 3: (2, 3), 4: (0, 0), 5: (0, 2), 10: (0, 0), 11: (0, 2),
                                  22: (0, 1), 23: (1, 2)}


def test_hand_multipoles_h2():
    # equal masses m = 1/4.  Leaves: COM = the particle, r = 0, moments (m, 0...).
    # box 7 holds p3 (.5,.75,.25) and p0 (.5,.5,.5): COM (.5,.625,.375); offsets +-(0,.125,-.125), so
    #   M0 = .5, M_yy = M_zz = 2 m .125^2 = .0078125, M_yz = -.0078125, every other second and every
    #   third moment 0; r = |(0,.125,.125)| = sqrt(2)/8 (children and particle box agree).
    # box 3 holds box 6 (p2 = (.75,.25,.5), N 1) and box 7 (N 2): COM (7/12, 1/2, 5/12);
    #   r = min(max(sqrt(14)/12, sqrt(14)/24 + sqrt(2)/8), sqrt(17)/12) = sqrt(14)/24 + sqrt(2)/8.
    # root: COM (.5, .5, .4375) (the mean of the four particles);
    #   r = min(max(|(-.25,0,.0625)|, sqrt(17)/48 + r_3), |(.25,.25,.1875)|) = sqrt(0.16015625).
    s = solver_for(H2, ppn=1)
    ids, idx = by_boxid(s)
    X = s.export("node_x") * 2.0 ** -32 + 0.5
    r = s.export("node_r")
    M = s.export("node_m")
    np.testing.assert_allclose(X[idx[7]], [0.5, 0.625, 0.375], rtol=0, atol=1e-9)
    np.testing.assert_allclose(X[idx[3]], [7 / 12, 0.5, 5 / 12], rtol=0, atol=1e-9)
    np.testing.assert_allclose(X[idx[1]], [0.5, 0.5, 0.4375], rtol=0, atol=1e-9)
    r7 = np.float32(np.sqrt(2) / 8)
    r3 = np.float32(np.sqrt(14) / 24 + float(r7))
    np.testing.assert_allclose(r[idx[7]], r7, rtol=1e-6)
    np.testing.assert_allclose(r[idx[3]], r3, rtol=1e-6)
    np.testing.assert_allclose(r[idx[1]], np.sqrt(0.16015625), rtol=1e-6)
    for leaf in (2, 6, 14, 15):
        assert r[idx[leaf]] == 0.0
        np.testing.assert_allclose(M[idx[leaf]], [0.25] + [0.0] * 16, atol=1e-12)
    # slots 1..6 = xx xy xz yy yz zz, 7..16 = the ten third moments
    want = np.zeros(17)
    want[0] = 0.5
    want[4] = want[6] = 0.0078125
    want[5] = -0.0078125
    np.testing.assert_allclose(M[idx[7]], want, rtol=0, atol=1e-9)


# hand walk of H2 (open boundary, h = 1e-3); leaves A = box 2 (p1), C = 6 (p2), E = 14 (p3),
# F = 15 (p0), interior B = 3, D = 7.  Separations: |A-B| = sqrt(17)/12 = .3436, |A-C| = .5590,
# |A-D| = .3062, |C-D| = .4677, |A-E| = .4330, |A-F| = .25, |E-F| = .3536; radii B .3327, D .1768.
HAND_LISTS = {
    # theta 0.7 (1/theta = 1.4286): A-B near (.3436 < .4753): at A the drain takes C (r 0: far) and
    # D (.3062 > .2525: far) as PC; at B, A is opened and carried down; C sees A far (CP) and D far
    # (CC); D sees A far (CP) and C far (CC); E and F see each other far.
    0.7: {"cc_d": {6: [7], 7: [6], 14: [15], 15: [14]},
          "cp_d": {6: [(0, 1)], 7: [(0, 1)]},
          "pc_d": {2: [6, 7]},
          "pp_d": {2: [(0, 1)], 6: [(1, 2)], 14: [(2, 3)], 15: [(3, 4)]}},
    # theta 0.5 (1/theta = 2): A-D is near (.3062 < .3536), so A's drain opens D into E and F (PC),
    # and D carries the opened A down to E and F, which see it far (CP).
    0.5: {"cc_d": {6: [7], 7: [6], 14: [15], 15: [14]},
          "cp_d": {6: [(0, 1)], 14: [(0, 1)], 15: [(0, 1)]},
          "pc_d": {2: [6, 14, 15]},
          "pp_d": {2: [(0, 1)], 6: [(1, 2)], 14: [(2, 3)], 15: [(3, 4)]}},
}


@pytest.mark.parametrize("theta", [0.7, 0.5])
def test_hand_lists_h2(theta):
    s = solver_for(H2, theta=theta, ppn=1, h=1e-3)
    ids, idx = by_boxid(s)
    want = HAND_LISTS[theta]
    for name in ("cc_d", "cp_d", "pc_d", "pp_d"):
        p, v = s.export(name + "_ptr"), s.export(name)
        got = {}
        for node in range(len(ids)):
            seg = v[p[node]:p[node + 1]]
            if not len(seg):
                continue
            if name in ("cp_d", "pp_d"):
                got[ids[node]] = [tuple(int(t) for t in q) for q in seg.reshape(-1, 2)]
            else:
                got[ids[node]] = [ids[int(c)] for c in seg]
        assert got == want[name], name


# ---- 2. the MAC verifier on host-built lists ----------------------------------------------------
WORKLOADS = {
    # name: (generator, periodic, theta, leaf)
    "uniform32k_t07_leaf64": (lambda ics: ics.uniform(32768, 1), True, 0.7, 64),
    "zeldovich24_t03_leaf32": (lambda ics: ics.zeldovich(24, 11), True, 0.3, 32),
    "plummer_open_t05_leaf16": (lambda ics: ics.plummer_spheres(20000, 12), False, 0.5, 16),
    "plummer_periodic_t07_leaf8": (lambda ics: ics.plummer_spheres(6000, 13), True, 0.7, 8),
}


def host_lists(x, periodic, theta, leaf):
    from tigergrav_amd import ics
    from tigergrav_amd.solver import LIST_NAMES, Solver

    n = x.shape[0]
    s = Solver(ewald=periodic, soft_len=ics.soft_len(n), m=1.0 / n, theta=theta, parts_per_node=leaf, device=-1)
    s.set_particles(x)
    T = s.tree_arrays()
    for k in LIST_NAMES:
        T[k] = s.export(k)
        T[k + "_ptr"] = s.export(k + "_ptr")
    return T, ics.soft_len(n)


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_mac_verifier_host_lists(name):
    from tigergrav_amd import ics

    gen, periodic, theta, leaf = WORKLOADS[name]
    x = gen(ics)
    T, h = host_lists(x, periodic, theta, leaf)
    bad = MacVerifier(T, theta, h, periodic).check()
    assert all(v == 0 for v in bad.values()), bad
    kinds = ("d", "e") if periodic else ("d",)
    for k in kinds:
        assert coverage_exact(T, T["leaves"], k) == []


def test_mac_verifier_rejects_other_theta_and_softening():
    """Lists of theta 0.7 checked as if they were of theta 0.65 / 0.75, or with a softening 100x
    larger, must show violations: the verifier tells these MACs apart."""
    from tigergrav_amd import ics

    x = ics.uniform(32768, 1)
    T, h = host_lists(x, True, 0.7, 64)
    for theta, hh in ((0.65, h), (0.75, h)):
        bad = MacVerifier(T, theta, hh, True).check()
        assert sum(bad.values()) > 0, (theta, hh)
    x = ics.plummer_spheres(20000, 12)
    T, h = host_lists(x, False, 0.5, 16)
    assert sum(MacVerifier(T, 0.5, 100 * h, False).check().values()) > 0
    assert sum(MacVerifier(T, 0.5, h, True).check().values()) > 0  # open lists are not periodic ones


def test_coverage_exact_detects_a_dropped_entry():
    from tigergrav_amd import ics

    x = ics.uniform(4096, 3)
    T, h = host_lists(x, True, 0.7, 32)
    lf = int(T["leaves"][5])
    assert coverage_exact(T, [lf], "d") == []
    p = T["pp_d_ptr"].copy()
    v = T["pp_d"].copy()
    b, e = p[lf], p[lf + 1]
    assert e - b >= 2
    T2 = dict(T)
    T2["pp_d"] = np.concatenate([v[:b], v[b + 2:]])
    T2["pp_d_ptr"] = np.concatenate([p[:lf + 1], p[lf + 1:] - 2])
    assert coverage_exact(T2, [lf], "d") == [lf]
