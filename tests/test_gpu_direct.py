"""The FP64 direct-summation accuracy reference (include/tg_direct.h) against the oracle.

The oracle's tgo_direct64 evaluates the same potential another way: per pair, with the Ewald sum
split at alpha = 2 and truncated at |n|^2 <= 25, |h|^2 <= 16 (both converged far below FP64
rounding), while the GPU splits at alpha = 6 and sums the Fourier part through structure
factors.  Agreement to ~1e-10 therefore checks the split, the images, the structure factors and
the constants (the oracle returns the reference's rounded 2.8372975 for the r -> 0 correction,
the GPU the exact limit: the sinks here have no coincident partner but themselves, whose share
of phi is m * 5e-8).
"""
import numpy as np
import pytest

import oracle_bind as ob
from tigergrav_amd import ics

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(5)
    yield "uniform_periodic", ics.uniform(2048, seed=21), True, None
    yield "uniform_open", ics.uniform(2048, seed=22), False, None
    yield "plummer_periodic", ics.plummer_spheres(4096, seed=23), True, None
    yield "masses_periodic", ics.uniform(1024, seed=24), True, rng.uniform(0.5, 1.5, 1024).astype(np.float32)


@pytest.mark.parametrize("name,x,ewald,mass", list(_cases()), ids=[c[0] for c in _cases()])
def test_direct_matches_oracle(name, x, ewald, mass):
    from tigergrav_amd.accuracy import direct_forces

    n = x.shape[0]
    h = ics.soft_len(n)
    m = 1.0 / n
    sinks = np.random.default_rng(len(name)).choice(n, 96, replace=False)
    print(f"direct {name} n={n} ewald={ewald}", flush=True)
    f = direct_forces(x, sinks, ewald, h, m, mass=mass)
    mj = (mass.astype(np.float64) if mass is not None else np.full(n, m))
    phi, g = ob.direct64(ewald, h, x[sinks], x, mj)
    dg = np.linalg.norm(f[:, 1:] - g, axis=1) / np.linalg.norm(g, axis=1)
    assert np.sqrt(np.mean(dg ** 2)) < 1e-10 and dg.max() < 1e-9, (np.sqrt(np.mean(dg ** 2)), dg.max())
    # the self pair: oracle constant vs exact limit, m * |2.8372975 - psi(0)| < m * 1e-7
    tol = 1e-7 * (mj[sinks] if ewald else 0.0) + 1e-10 * np.abs(phi)
    assert np.all(np.abs(f[:, 0] - phi) <= tol), np.max(np.abs(f[:, 0] - phi) - tol)


def test_direct_two_body_closed_form():
    """Open boundaries, two bodies beyond the softening length: Newton exactly."""
    from tigergrav_amd.accuracy import direct_forces

    x = ics.double_to_pos(np.array([[0.3, 0.5, 0.5], [0.7, 0.5, 0.5]]))
    h = 0.01
    f = direct_forces(x, [0, 1], False, h, 0.5)
    r = (float(x[1, 0]) - float(x[0, 0])) * 2.0 ** -32  # the separation after fixed-point rounding
    self_phi = -0.5 * (14.0 / 5.0) / h
    assert f[0, 1] == pytest.approx(0.5 / r ** 2, rel=1e-12)
    assert f[1, 1] == pytest.approx(-0.5 / r ** 2, rel=1e-12)
    assert f[0, 0] == pytest.approx(-0.5 / r + self_phi, rel=1e-12)
    assert abs(f[0, 2]) < 1e-12 and abs(f[0, 3]) < 1e-12


def test_direct_periodic_symmetry():
    """Periodic: a pair half a box apart along x feels no net force (the images cancel), and a
    lattice of 2^3 points on a half-box grid feels none at all."""
    from tigergrav_amd.accuracy import direct_forces

    x = ics.double_to_pos(np.array([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5]]))
    f = direct_forces(x, [0, 1], True, 1e-3, 0.5)
    assert np.all(np.abs(f[:, 1:]) < 1e-10)
    g = np.stack(np.meshgrid([0.1, 0.6], [0.2, 0.7], [0.3, 0.8], indexing="ij"), -1).reshape(-1, 3)
    f = direct_forces(ics.double_to_pos(g), np.arange(8), True, 1e-3, 1.0 / 8)
    assert np.all(np.abs(f[:, 1:]) < 1e-9)


@pytest.mark.parametrize("ewald", [True, False], ids=["periodic", "open"])
def test_direct_gradient_matches_fp64_finite_differences(ewald):
    """tg_direct_gradient against central differences of tg_direct_forces (both FP64): every
    particle a sink, L = sum_i w_i . g_i, one coordinate or mass moved at a time."""
    from tigergrav_amd.accuracy import direct_forces, direct_gradient

    n = 192
    x = ics.plummer_spheres(n, 91)
    h, m = 4.0 * ics.soft_len(n), 1.0 / n  # larger softening: some pairs inside the spline
    rng = np.random.default_rng(92)
    w = rng.standard_normal((n, 3)).astype(np.float32)
    mass = (m * rng.uniform(0.5, 1.5, n)).astype(np.float32)
    allk = np.arange(n)

    def loss(xx, mm):
        return float(np.sum(w.astype(np.float64) * direct_forces(xx, allk, ewald, h, m, mass=mm)[:, 1:]))

    pick = rng.choice(n, 6, replace=False)
    g = direct_gradient(x, w, pick, ewald, h, m, mass=mass)
    step = 2 ** 12  # 2^-20 of the box
    print(f"direct gradient ewald={ewald}", flush=True)
    for r, k in enumerate(pick):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, d] += step
            xm[k, d] -= step
            fd = (loss(xp, mass) - loss(xm, mass)) / (2 * step * 2.0 ** -32)
            assert abs(fd - g[r, d]) <= 1e-6 * abs(g[r, d]) + 1e-9 * np.abs(g[:, :3]).max(), (k, d, fd, g[r, d])
        mp, mm_ = mass.copy(), mass.copy()
        dm = np.float32(2.0 ** -20)
        mp[k] += dm
        mm_[k] -= dm
        fd = (loss(x, mp) - loss(x, mm_)) / (float(mp[k]) - float(mm_[k]))
        assert abs(fd - g[r, 3]) <= 1e-7 * abs(g[r, 3]) + 1e-9 * np.abs(g[:, 3]).max(), (k, fd, g[r, 3])
