"""The tabulated periodic correction (tigergrav_amd/csrc/tg_ewald_table.cpp, tg_math.hpp green_ewald).

* The FP64 generator tg_ewald_green64 against the oracle's FP64 green_ewald (the reference's
  truncation, tigergrav/green.hpp:189-286) on orders 0..4, and its orders 5..7 against central
  differences of the order below.
* Symmetry: permuting coordinates permutes the components; flipping a coordinate's sign flips the
  components with an odd count of that index.
* The kernels' interpolation (|x| sorted into the wedge u0 >= u1 >= u2, third-order Taylor about
  the nearest point of a 1/64 grid of float-rounded values there, components permuted back, signs
  by octant), restated here in FP64, against the generator at random separations: the error
  bound the kernels document.  The kernels themselves are checked against
  the oracle by the GPU kernel tests.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import oracle_bind as ob
from tigergrav_amd._lib import lib

ORDER = 7
SLOTS = [(n0, n1, n - n0 - n1) for n in range(ORDER + 1) for n0 in range(n, -1, -1) for n1 in range(n - n0, -1, -1)]
SIDX = {s: i for i, s in enumerate(SLOTS)}
ORD = np.array([sum(s) for s in SLOTS])
NS = len(SLOTS)


def g64(x):
    L = lib()
    L.tg_ewald_green64.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    D = np.zeros((len(x), NS))
    assert L.tg_ewald_green64(x.ctypes.data, len(x), D.ctypes.data) == 0
    return D


def test_slot_count():
    assert NS == 120 and SLOTS[:35][20] == (4, 0, 0) and SLOTS[34] == (0, 0, 4)


def test_generator_matches_oracle_orders_0_to_4():
    rng = np.random.default_rng(11)
    X = rng.uniform(-0.5, 0.5, (200, 3))
    X[0] = (0.5, 0.0, 0.0)
    X[1] = (0.25, -0.5, 0.5)
    X[2] = (1e-3, 2e-3, -1e-3)
    D = g64(X)
    O = np.array([ob.green_ewald(x, fp64=True) for x in X])
    for k in range(5):
        m = ORD[:35] == k
        err = np.max(np.abs(D[:, :35][:, m] - O[:, m])) / np.max(np.abs(O[:, m]))
        assert err < 1e-9, (k, err)


def test_generator_high_orders_are_derivatives():
    rng = np.random.default_rng(12)
    X = rng.uniform(-0.45, 0.45, (30, 3))
    D = g64(X)
    h = 1e-4
    for k in (5, 6, 7):
        scale = np.max(np.abs(D[:, ORD == k]))
        for s in np.nonzero(ORD == k)[0]:
            c = SLOTS[s]
            a = next(q for q in range(3) if c[q] > 0)
            lo = list(c)
            lo[a] -= 1
            e = np.zeros(3)
            e[a] = h
            fd = (g64(X + e)[:, SIDX[tuple(lo)]] - g64(X - e)[:, SIDX[tuple(lo)]]) / (2 * h)
            assert np.max(np.abs(fd - D[:, s])) < 2e-6 * scale, (k, c)


def test_generator_symmetries():
    rng = np.random.default_rng(13)
    X = rng.uniform(-0.5, 0.5, (40, 3))
    D = g64(X)
    for pm in itertools.permutations(range(3)):
        Y = X[:, pm]  # y_a = x_(pm[a])
        DY = g64(Y)
        for k, c in enumerate(SLOTS):
            m = [0, 0, 0]
            for a in range(3):
                m[pm[a]] = c[a]
            np.testing.assert_allclose(DY[:, k], D[:, SIDX[tuple(m)]], rtol=1e-10, atol=1e-10 * np.max(np.abs(D[:, ORD == sum(c)])))
    for a in range(3):
        Y = X.copy()
        Y[:, a] = -Y[:, a]
        DY = g64(Y)
        sign = np.array([-1.0 if c[a] % 2 else 1.0 for c in SLOTS])
        for k in range(ORDER + 1):
            m = ORD == k
            np.testing.assert_allclose(DY[:, m], (D * sign)[:, m], rtol=1e-10, atol=1e-10 * np.max(np.abs(D[:, m])))


def interpolate(x, inv=64, npts=33):
    """The kernels' scheme in FP64 from float-rounded table values: |x| sorted into the wedge
    u0 >= u1 >= u2 by the compare-exchanges (0,1), (1,2), (0,1); the grid point's slots that hold
    the third axis at most once from the table, the others from the trace relation (the Laplacian
    is -4 pi); a third-order Taylor expansion about the point for the components that hold the
    third axis at most once, the others again from the trace relation; the exchanges undone on the
    components in reverse order; signs by octant."""
    u = np.abs(x).copy()
    sg = np.where(x < 0, -1.0, 1.0)
    swaps = []
    for a, b in ((0, 1), (1, 2), (0, 1)):
        c = u[:, b] > u[:, a]
        u[c, a], u[c, b] = u[c, b], u[c, a].copy()
        swaps.append((a, b, c))
    gi = np.minimum(np.floor(u * inv + 0.5), npts - 1)
    assert np.all(gi[:, 0] >= gi[:, 1]) and np.all(gi[:, 1] >= gi[:, 2])
    d = u - gi / inv
    T = g64(gi / inv).astype(np.float32).astype(np.float64)
    # the table stores the slots holding the third axis at most once; the others come from the
    # trace relation, in slot order
    for k, c in enumerate(SLOTS):
        if c[2] >= 2:
            lap = -4.0 * math.pi if c == (0, 0, 2) else 0.0
            T[:, k] = lap - T[:, SIDX[(c[0] + 2, c[1], c[2] - 2)]] - T[:, SIDX[(c[0], c[1] + 2, c[2] - 2)]]
    out = np.zeros((len(x), 35))
    mono = [j for j in range(1, 20)]  # slots of orders 1..3
    for k in range(35):
        c = SLOTS[k]
        if c[2] >= 2:
            lap = -4.0 * math.pi if c == (0, 0, 2) else 0.0
            out[:, k] = lap - out[:, SIDX[(c[0] + 2, c[1], c[2] - 2)]] - out[:, SIDX[(c[0], c[1] + 2, c[2] - 2)]]
            continue
        v = T[:, k].copy()
        for j in mono:
            J = SLOTS[j]
            coef = 1.0 / (math.factorial(J[0]) * math.factorial(J[1]) * math.factorial(J[2]))
            v += T[:, SIDX[(c[0] + J[0], c[1] + J[1], c[2] + J[2])]] * coef * d[:, 0] ** J[0] * d[:, 1] ** J[1] * d[:, 2] ** J[2]
        out[:, k] = v
    for a, b, c in reversed(swaps):
        for k in range(35):
            n = list(SLOTS[k])
            n[a], n[b] = n[b], n[a]
            p = SIDX[tuple(n)]
            if p > k:
                out[c, k], out[c, p] = out[c, p], out[c, k].copy()
    for k in range(35):
        out[:, k] *= np.prod(sg ** np.array(SLOTS[k]), axis=1)
    return out


def test_constant_laplacian():
    """The correction's Laplacian is -4 pi over the cell (the periodic potential's uniform
    background; the nearest image's own Laplacian is the point mass), so every higher derivative
    tensor is traceless: D_aa = -4 pi and sum_a D_{k+2a} = 0 for |k| = 1..5 (the relation the
    kernels use for the components that hold the wedge's third axis twice or more), to 1e-8 of
    each order's scale."""
    rng = np.random.default_rng(21)
    X = rng.uniform(-0.5, 0.5, (300, 3))
    X = X[np.linalg.norm(X, axis=1) > 0.05]
    D = g64(X)
    np.testing.assert_allclose(D[:, SIDX[(2, 0, 0)]] + D[:, SIDX[(0, 2, 0)]] + D[:, SIDX[(0, 0, 2)]], -4.0 * math.pi,
                               rtol=1e-9)
    for k, c in enumerate(SLOTS):
        if 1 <= sum(c) <= ORDER - 2:
            tr = sum(D[:, SIDX[(c[0] + 2 * (a == 0), c[1] + 2 * (a == 1), c[2] + 2 * (a == 2))]] for a in range(3))
            scale = np.max(np.abs(D[:, ORD == sum(c) + 2]))
            assert np.max(np.abs(tr)) < 1e-8 * scale, (c, np.max(np.abs(tr)), scale)


@pytest.mark.parametrize("region", ["cell", "near_origin", "near_corner"])
def test_interpolation_error_bound(region):
    rng = np.random.default_rng({"cell": 14, "near_origin": 15, "near_corner": 16}[region])
    if region == "cell":
        X = rng.uniform(-0.5, 0.5, (400, 3))
    elif region == "near_origin":
        X = rng.uniform(-0.03, 0.03, (200, 3))
    else:
        X = np.sign(rng.uniform(-1, 1, (200, 3))) * rng.uniform(0.47, 0.5, (200, 3))
    A = interpolate(X)
    E = g64(X)[:, :35]
    scale_all = np.max(np.abs(g64(rng.uniform(-0.5, 0.5, (200, 3)))[:, :35]), axis=0)
    for k in range(5):
        m = ORD[:35] == k
        scale = np.max(scale_all[m])
        err = np.max(np.abs(A[:, m] - E[:, m])) / scale
        assert err < 5e-6, (region, k, err)
