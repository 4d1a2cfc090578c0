"""Pin the oracle against mathematics, independently of the reference.

Each check is chosen so a plausible slip (a dropped term, a sign, a wrong index, a swapped
operand) fails it:
  * Newton's law beyond the softening length, the spline's r -> 0 values, the phi_self offset;
  * green_direct against finite differences of -1/r;
  * the Ewald correction against an independent Ewald split (alpha = 4, numpy, FP64) --
    the correction must not depend on the split parameter --, its r -> 0 constant, its
    symmetry zeros, antisymmetry and periodicity;
  * M2P / M2L / P2L against brute-force sums over the particles of a well-separated cluster, with
    the truncation error shrinking like (r/d)^4;
  * the reference's float erfc / exp / sin approximations against libm within their published
    error bounds;
  * the float restatement (tgo_) against the FP64 restatement (tgo64_) on all seeded cases.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from kernel_cases import PARAMS, double_to_pos, kernel_cases, pos_to_double, p2m

S = 2.0 ** 32


def sym_lists():
    m2 = [(a, b) for a in range(3) for b in range(a, 3)]
    m3 = [(a, b, c) for a in range(3) for b in range(a, 3) for c in range(b, 3)]
    m4 = [(a, b, c, d) for a in range(3) for b in range(a, 3) for c in range(b, 3) for d in range(c, 3)]
    return m2, m3, m4


M2, M3, M4 = sym_lists()


def slot(idx):
    idx = tuple(sorted(idx))
    if len(idx) == 0:
        return 0
    if len(idx) == 1:
        return 1 + idx[0]
    if len(idx) == 2:
        return 4 + M2.index(idx)
    if len(idx) == 3:
        return 10 + M3.index(idx)
    return 20 + M4.index(idx)


# ----------------------------------------------------------------------------------------------
def test_pp_newton_and_spline_limits(oracle):
    P = dict(PARAMS["periodic"])
    h, m = P["soft_len"], P["m"]
    X = double_to_pos(np.array([[0.5, 0.5, 0.5]]))
    for r in (3 * h, 10 * h, 0.1):
        Y = X.copy()
        Y[0, 1] -= int(round(r * S))
        rr = int(round(r * S)) / S
        out, _ = ob.call_kernel("tgo_", "PP_direct", P, X, Y)
        phi_self = (-315.0 / 128.0) * m / h
        assert out[0, 2] == pytest.approx(-m / rr ** 2, rel=2e-6)
        assert abs(out[0, 1]) < 1e-12 and abs(out[0, 3]) < 1e-12
        assert out[0, 0] == pytest.approx(-m / rr - phi_self, rel=2e-6)
    # coincident pair: spline at r=0, force zero, phi = -m 14/5 / h - phi_self  (quirk Q2)
    out, _ = ob.call_kernel("tgo_", "PP_direct", P, X, X)
    assert np.all(out[0, 1:] == 0)
    assert out[0, 0] == pytest.approx(-m * 14.0 / 5.0 / h + 315.0 / 128.0 * m / h, rel=1e-6)


def test_green_direct_finite_difference(oracle):
    rng = np.random.default_rng(3)
    for _ in range(5):
        x = rng.normal(size=3) * 0.1
        D = ob.green_direct(x, fp64=True)
        Df = ob.green_direct(x.astype(np.float32)).astype(np.float64)
        np.testing.assert_allclose(Df, D, rtol=2e-5, atol=1e-6 * np.abs(D).max())
        f = lambda y: -1.0 / np.linalg.norm(y)
        assert D[0] == pytest.approx(f(x), rel=1e-14)
        eps = 1e-5
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            # each order's tensor is the gradient of the previous order's
            Dp, Dm = ob.green_direct(x + e, fp64=True), ob.green_direct(x - e, fp64=True)
            num = (Dp - Dm) / (2 * eps)
            assert num[0] == pytest.approx(D[slot((a,))], rel=1e-6)
            for i, idx in enumerate(M2):
                assert num[4 + i] == pytest.approx(D[slot(idx + (a,))], rel=1e-5, abs=1e-8 * abs(D).max())
            for i, idx in enumerate(M3):
                assert num[10 + i] == pytest.approx(D[slot(idx + (a,))], rel=1e-5, abs=1e-8 * abs(D).max())


def ewald_psi_alpha(x, alpha, nmax=6, hmax=14):
    """Periodic potential correction (relative to -1/|x|) of a unit mass, FP64, with Ewald
    split parameter alpha (independent of the alpha = 2 split the reference uses)."""
    from scipy.special import erfc

    r0 = np.linalg.norm(x)
    n = np.array([(i, j, k) for i in range(-nmax, nmax + 1) for j in range(-nmax, nmax + 1) for k in range(-nmax, nmax + 1)],
                 np.float64)
    d = np.linalg.norm(x[None, :] - n, axis=1)
    real = -np.sum(erfc(alpha * d) / d)
    hv = np.array([(i, j, k) for i in range(-hmax, hmax + 1) for j in range(-hmax, hmax + 1) for k in range(-hmax, hmax + 1)
                   if (i, j, k) != (0, 0, 0)], np.float64)
    h2 = np.sum(hv * hv, axis=1)
    four = -np.sum(np.exp(-math.pi ** 2 * h2 / alpha ** 2) / h2 * np.cos(2 * math.pi * hv @ x)) / math.pi
    return real + four + math.pi / alpha ** 2 + 1.0 / r0


def test_ewald_correction_split_independent(oracle):
    rng = np.random.default_rng(5)
    for _ in range(4):
        x = rng.uniform(-0.5, 0.5, 3)
        D = ob.green_ewald(x, fp64=True)
        assert D[0] == pytest.approx(ewald_psi_alpha(x, 4.0), abs=2e-9)
        # gradient of the potential = the first-order tensor
        eps = 1e-5
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            num = (ob.green_ewald(x + e, fp64=True)[0] - ob.green_ewald(x - e, fp64=True)[0]) / (2 * eps)
            assert num == pytest.approx(D[1 + a], abs=1e-7)


def test_ewald_correction_symmetry_and_limits(oracle):
    # r -> 0 value is the lattice constant 2.8372975 and the r > 0 formula approaches it
    D0 = ob.green_ewald(np.array([0.0, 0.0, 0.0]), fp64=True)
    assert D0[0] == 2.8372975 and np.all(D0[1:] == 0)
    Dsmall = ob.green_ewald(np.array([1e-4, 0.0, 0.0]), fp64=True)
    assert Dsmall[0] == pytest.approx(2.8372975, abs=1e-6)
    # zero correction force at the half-box symmetry points
    for p in ([0.5, 0, 0], [0.5, 0.5, 0], [0.5, 0.5, 0.5]):
        D = ob.green_ewald(np.array(p, np.float64), fp64=True)
        # at x = 1/2 e the total periodic force vanishes: correction grad = -grad(-1/r) of the
        # nearest image => D1 = -x/|x|^3 ... check the periodic total instead
        g_total = D[1:4] + np.array(p) / np.linalg.norm(p) ** 3
        np.testing.assert_allclose(g_total, 0.0, atol=1e-9)
    rng = np.random.default_rng(9)
    for _ in range(3):
        x = rng.uniform(-0.45, 0.45, 3)
        Dp, Dm = ob.green_ewald(x, fp64=True), ob.green_ewald(-x, fp64=True)
        odd = np.array([1] + [-1] * 3 + [1] * 6 + [-1] * 10 + [1] * 15, np.float64)
        np.testing.assert_allclose(Dm, odd * Dp, rtol=1e-9, atol=1e-9)
    # periodicity of the full periodic potential (correction + nearest-image -1/r)
    x = np.array([0.45, -0.2, 0.1])
    xs = x - np.array([1.0, 0.0, 0.0])
    assert ewald_psi_alpha(x, 4.0) - 1.0 / np.linalg.norm(x) == pytest.approx(
        ewald_psi_alpha(xs, 4.0) - 1.0 / np.linalg.norm(xs), abs=1e-9)


def test_reference_math_approximations(oracle):
    L = ob.lib()
    import ctypes as C

    from scipy.special import erfc

    for x in np.linspace(0.0, 7.2, 200):
        e = C.c_float()
        v = L.tgo_erfcexp(float(x), C.byref(e))
        assert abs(v - erfc(x)) < 6e-7  # A&S 7.1.26: |error| <= 1.5e-7, plus float rounding of ~1.4-sized terms
        xf = np.float32(x)
        arg = float(np.float32(-xf) * xf)  # the argument as the float code forms it
        # 2^(arg/ln2): the float product arg*(1/ln2) carries |arg/ln2| ulps of exponent error
        assert e.value == pytest.approx(math.exp(arg), rel=1.2e-7 * (abs(arg) * 1.4427 + 8), abs=1e-30)
    for x in np.linspace(-40.0, 40.0, 400):
        assert abs(L.tgo_sin(float(x)) - math.sin(float(np.float32(x)))) < 2e-6
    for x in np.linspace(-80.0, 80.0, 161):
        assert L.tgo_exp(float(x)) == pytest.approx(math.exp(float(np.float32(x))), rel=1.2e-7 * (abs(x) * 1.4427 + 8))


def brute_sum(xs_d, src_d, m):
    d = xs_d[:, None, :] - src_d[None, :, :]
    r = np.linalg.norm(d, axis=2)
    g = -m * np.sum(d / r[:, :, None] ** 3, axis=1)
    phi = -m * np.sum(1.0 / r, axis=1)
    return phi, g


def test_m2p_truncation_order(oracle):
    """M2P force error against the brute-force cluster sum falls like (r/d)^4 (octupole kept)."""
    P = dict(PARAMS["open"])
    m = P["m"]
    rng = np.random.default_rng(11)
    base = 0.5 + 0.02 * (rng.random((40, 3)) - 0.5)
    sinks = np.array([[0.5, 0.5, 0.5]]) + np.array([[0.0, 0.0, 0.0]])
    errs = []
    for dist in (0.1, 0.2, 0.4):
        src = base + np.array([dist, 0.0, 0.0])
        com = double_to_pos(src.mean(axis=0))
        src_q = pos_to_double(double_to_pos(src))
        M = np.zeros(1, ob.MULTI_DTYPE)
        M[0]["m"] = p2m(src_q, m, com)
        M[0]["x"] = com
        xs = double_to_pos(sinks)
        out, _ = ob.call_kernel("tgo64_", "PC_direct", P, xs, M)
        _, g = brute_sum(pos_to_double(xs), src_q, m)
        errs.append(np.linalg.norm(out[0, 1:] - g[0]) / np.linalg.norm(g[0]))
    rate = math.log(errs[0] / errs[2]) / math.log(4.0)
    assert rate > 3.6, errs


def l2p(L, dx):
    """Evaluate a 4th-order local expansion (potential + acceleration) -- Taylor series."""
    phi = L[0]
    g = -np.array([L[1], L[2], L[3]], np.float64)
    for order, lst in ((1, [(i,) for i in range(3)]), (2, M2), (3, M3), (4, M4)):
        for idx in lst:
            mult = math.factorial(order)
            for v in set(idx):
                mult //= math.factorial(idx.count(v))
            c = L[slot(idx)] * mult / math.factorial(order)
            phi += c * np.prod([dx[i] for i in idx])
    # gradient by the same series (one order lower)
    for a in range(3):
        for order, lst in ((1, [(i,) for i in range(3)]), (2, M2), (3, M3)):
            for idx in lst:
                mult = math.factorial(order)
                for v in set(idx):
                    mult //= math.factorial(idx.count(v))
                c = L[slot(idx + (a,))] * mult / math.factorial(order)
                g[a] -= c * np.prod([dx[i] for i in idx])
    return phi, g


@pytest.mark.parametrize("kind", ["CC_direct", "CP_direct"])
def test_m2l_p2l_against_brute_force(oracle, kind):
    P = dict(PARAMS["open"])
    m = P["m"]
    rng = np.random.default_rng(17)
    src = 0.5 + 0.02 * (rng.random((30, 3)) - 0.5) + np.array([0.25, 0.05, -0.03])
    src_q = pos_to_double(double_to_pos(src))
    X = double_to_pos(np.array([0.5, 0.5, 0.5]))
    if kind == "CC_direct":
        com = double_to_pos(src_q.mean(axis=0))
        M = np.zeros(1, ob.MULTI_DTYPE)
        M[0]["m"] = p2m(src_q, m, com)
        M[0]["x"] = com
        L, _ = ob.call_kernel("tgo64_", kind, P, X, M)
    else:
        L, _ = ob.call_kernel("tgo64_", kind, P, X, double_to_pos(src_q))
    pts = 0.5 + 0.01 * (rng.random((5, 3)) - 0.5)
    phi_b, g_b = brute_sum(pts, src_q, m)
    for i, p in enumerate(pts):
        phi, g = l2p(L, p - pos_to_double(X))
        assert phi == pytest.approx(phi_b[i], rel=2e-5)
        np.testing.assert_allclose(g, g_b[i], rtol=2e-4 if kind == "CC_direct" else 1e-5)


DIRECT_OF = {"PP_ewald": "PP_direct", "PC_ewald": "PC_direct", "CC_ewald": "CC_direct", "CP_ewald": "CP_direct"}


@pytest.mark.parametrize("cfg", ["periodic", "open"])
@pytest.mark.parametrize("kind", ob.KINDS)
def test_float_restatement_close_to_fp64(oracle, cfg, kind):
    """tgo_ (reference float arithmetic) vs tgo64_ (same formula in FP64), rms relative.

    The Ewald kernels return a periodic *correction* that the reference forms as (lattice sum) -
    (Newtonian term) in float; for close pairs that cancels, so their error is measured against
    the scale of the total they are added to: correction + the matching direct-kernel term."""
    P = PARAMS[cfg]
    num = den = 0.0
    for c in kernel_cases(kind, P):
        if c["name"].endswith(("symmetry_points", "coincident")):
            continue  # limits checked exactly in the tests above
        sl = slice(1, 4) if kind[:2] in ("PP", "PC") else slice(0, 35)
        a = ob.call_kernel("tgo_", kind, P, c["x"], c["y"], do_phi=c["do_phi"])[0]
        b = ob.call_kernel("tgo64_", kind, P, c["x"], c["y"], do_phi=c["do_phi"])[0]
        a, b = a[..., sl], b[..., sl]
        num += np.sum((a - b) ** 2)
        den += np.sum(b ** 2)
        if kind in DIRECT_OF:
            Pd = dict(P, ewald=1)
            dterm = ob.call_kernel("tgo64_", DIRECT_OF[kind], Pd, c["x"], c["y"])[0][..., sl]
            den += np.sum(dterm ** 2)
    rel = math.sqrt(num / den)
    assert rel < 2e-6, rel
