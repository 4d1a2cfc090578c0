"""The adjoint of the batched solve (tg_solver_adjoint) against finite differences of the same
solve.

The adjoint differentiates the solve with its tree, interaction lists and expansion centres held
fixed -- exactly what tg_solver_update_particles re-evaluates.  For L = sum_i w_i . g_i and a random
direction d (every particle's position, or every mass, moved at once), the central difference
(L(x + d) - L(x - d)) / 2 of two frozen-tree solves must equal grad . d.  Moving every particle at
once makes the change of L large against the solve's single-precision rounding, whose floor in
these differences is about 1e-5 absolute (the tolerance carries it); the masses enter linearly, so
that check is tight.
"""
import numpy as np
import pytest

from tigergrav_amd import ics

pytestmark = pytest.mark.gpu

def _edge(x, k=64, seed=85):
    """k particles moved to within 2^12 position units of the periodic box edge, so that the
    directional steps carry some of them across it (the frozen tree must stay continuous there)."""
    x = x.copy()
    r = np.random.default_rng(seed)
    for d in range(3):
        sel = r.choice(x.shape[0], k, replace=False)
        x[sel, d] = (2 ** 31 - 2 ** 12 + r.integers(0, 2 ** 13, k)).astype(np.int64).astype(np.int32)
    return x


CASES = {
    # name: (generator, ewald, theta, leaf, position step in 2^-32 box units)
    "uniform_periodic": (lambda: ics.uniform(4096, 81), True, 0.7, 32, 2 ** 14),
    "plummer_open": (lambda: ics.plummer_spheres(8192, 82), False, 0.7, 64, 2 ** 14),
    "zeldovich_periodic": (lambda: ics.zeldovich(16, 83), True, 0.5, 64, 2 ** 18),  # near a lattice: small forces
    "box_edge_periodic": (lambda: _edge(ics.uniform(4096, 84)), True, 0.7, 32, 2 ** 14),
}


def _loss(s, w, x, mass=None):
    s.update_particles(x, mass)
    s.solve()
    s.synchronize()
    return float(np.sum(w * s.forces()[:, 1:]))


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_directional_derivatives(case):
    from tigergrav_amd.solver import Solver

    gen, ewald, theta, ppn, step_max = CASES[case]
    x = gen()
    n = x.shape[0]
    h, m = ics.soft_len(n), 1.0 / n
    print(f"adjoint {case}: N={n} ewald={ewald} theta={theta} leaf={ppn}", flush=True)
    s = Solver(ewald=ewald, soft_len=h, m=m, theta=theta, parts_per_node=ppn, device=0)
    s.set_particles(x)
    rng = np.random.default_rng(len(case))
    w = rng.standard_normal((n, 3)).astype(np.float32)
    s.solve()
    s.synchronize()
    grad, ms = s.adjoint(w, with_time=True)
    assert np.all(np.isfinite(grad))
    # positions: every particle moved by up to step_max (2^-18 .. 2^-14 of the box, inside every leaf)
    for k in range(3):
        step = rng.integers(-step_max, step_max, size=x.shape).astype(np.int64)
        xp = (x.astype(np.int64) + step).astype(np.int32)
        xm = (x.astype(np.int64) - step).astype(np.int32)
        fd = 0.5 * (_loss(s, w, xp) - _loss(s, w, xm))
        pred = float(np.sum(grad[:, :3] * step * 2.0 ** -32))
        print(f"  positions {k}: fd {fd:.6e} adjoint {pred:.6e} rel {abs(fd - pred) / abs(pred):.2e}", flush=True)
        assert abs(fd - pred) <= 3e-3 * abs(pred) + 2e-5
    # masses: L is linear in them
    for k in range(2):
        dm = (m * rng.uniform(-1e-2, 1e-2, n)).astype(np.float64)
        mp = (m + dm).astype(np.float32)
        mm = (m - dm).astype(np.float32)
        dmf = 0.5 * (mp.astype(np.float64) - mm.astype(np.float64))  # the steps the float masses carry
        fd = 0.5 * (_loss(s, w, x, mp) - _loss(s, w, x, mm))
        pred = float(np.sum(grad[:, 3] * dmf))
        print(f"  masses {k}: fd {fd:.6e} adjoint {pred:.6e} rel {abs(fd - pred) / abs(pred):.2e}", flush=True)
        assert abs(fd - pred) <= 1e-3 * abs(pred)
    print(f"  adjoint device time {ms:.3f} ms", flush=True)


@pytest.mark.parametrize("theta", [0.7, 0.3])
def test_adjoint_against_direct_sum_gradient(theta):
    """The solve's adjoint against the FP64 direct-sum gradient of the same loss: they differ by
    the tree's own approximation error, which shrinks with theta."""
    from tigergrav_amd.accuracy import direct_gradient
    from tigergrav_amd.solver import Solver

    n = 4096
    x = ics.plummer_spheres(n, 93)
    h, m = ics.soft_len(n), 1.0 / n
    rng = np.random.default_rng(94)
    w = rng.standard_normal((n, 3)).astype(np.float32)
    s = Solver(ewald=True, soft_len=h, m=m, theta=theta, parts_per_node=32, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    g = s.adjoint(w)
    pick = np.sort(rng.choice(n, 512, replace=False))
    gd = direct_gradient(x, w, pick, True, h, m)
    ex = np.linalg.norm(g[pick, :3] - gd[:, :3], axis=1) / np.linalg.norm(gd[:, :3], axis=1)
    em = np.abs(g[pick, 3] - gd[:, 3]) / np.abs(gd[:, 3])
    rx, rm = float(np.sqrt(np.mean(ex ** 2))), float(np.sqrt(np.mean(em ** 2)))
    print(f"adjoint vs direct gradient theta={theta}: rms rel d/dx {rx:.2e}, d/dm {rm:.2e}", flush=True)
    assert rx < (0.05 if theta > 0.5 else 0.01)


def test_torch_autograd_function():
    """The torch.autograd.Function routes backward through the adjoint: pos.grad and mass.grad of
    (w * acc).sum() equal the solver's adjoint, and a directional difference through the function
    itself (positions moved along a random direction, tree rebuilt each time) agrees with them."""
    import torch

    from tigergrav_amd.autograd import tree_accelerations
    from tigergrav_amd.solver import Solver

    n = 2048
    x = ics.uniform(n, 95)
    pos0 = (x.astype(np.float64) * 2.0 ** -32 + 0.5)
    h = ics.soft_len(n)
    rng = np.random.default_rng(96)
    w = torch.from_numpy(rng.standard_normal((n, 3)))
    pos = torch.tensor(pos0, requires_grad=True)
    mass = torch.full((n,), 1.0 / n, dtype=torch.float64, requires_grad=True)
    acc = tree_accelerations(pos, mass, ewald=True, soft_len=h, theta=0.7, leaf=32)
    (w * acc).sum().backward()
    s = Solver(ewald=True, soft_len=h, m=1.0 / n, theta=0.7, parts_per_node=32, device=0)
    s.set_particles(ics.double_to_pos(pos0), np.full(n, 1.0 / n, np.float32))
    s.solve()
    s.synchronize()
    g = s.adjoint(w.numpy())
    np.testing.assert_allclose(pos.grad.numpy(), g[:, :3], rtol=1e-6, atol=1e-9 * np.abs(g[:, :3]).max())
    np.testing.assert_allclose(mass.grad.numpy(), g[:, 3], rtol=1e-6, atol=1e-9 * np.abs(g[:, 3]).max())
    # directional check through the function (each call builds its own tree; the step keeps the lists)
    d = rng.uniform(-1, 1, (n, 3)) * 2.0 ** -20
    with torch.no_grad():
        lp = float((w * tree_accelerations(torch.tensor(pos0 + d), mass, True, h, 0.7, 32)).sum())
        lm = float((w * tree_accelerations(torch.tensor(pos0 - d), mass, True, h, 0.7, 32)).sum())
    pred = float(np.sum(pos.grad.numpy() * d))
    print(f"torch directional: fd {0.5 * (lp - lm):.6e} adjoint {pred:.6e}", flush=True)
    assert abs(0.5 * (lp - lm) - pred) <= 3e-3 * abs(pred) + 2e-5


def test_adjoint_rejects_a_leaf_beyond_64_particles():
    """A leaf at the tree's depth limit can hold more than 64 coincident particles; the adjoint,
    which takes a leaf's sinks in one 64-lane group, refuses it with an error instead of leaving
    the extra particles out."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = np.concatenate([np.repeat(ics.uniform(1, 61), 80, axis=0), ics.uniform(3000, 62)])
    n = x.shape[0]
    print(f"adjoint oversized leaf: N={n}", flush=True)
    s = Solver(ewald=True, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.7, parts_per_node=32, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    with pytest.raises(Exception, match="64 particles"):
        s.adjoint(np.ones((n, 3), np.float32))


FP64_CASES = {
    # name: (generator, ewald, theta, leaf)
    "uniform_periodic": (lambda: ics.uniform(1024, 181), True, 0.7, 16),
    "box_edge_periodic": (lambda: _edge(ics.uniform(1024, 184), k=16), True, 0.7, 16),
    "zeldovich_periodic": (lambda: ics.zeldovich(8, 183), True, 0.7, 8),
    "plummer_open": (lambda: ics.plummer_spheres(8192, 182), False, 0.7, 64),
}


@pytest.mark.parametrize("case", list(FP64_CASES))
def test_adjoint_against_fp64_frozen_tree(case):
    """The adjoint's directional derivatives against central differences of the oracle's FP64
    frozen-tree solve (tgo64_frozen_solve: the same tree, lists and expansion centres, moments
    recomputed in FP64 about them, the FP64 kernels and FP64 L2L/L2P) -- the function the adjoint
    differentiates, free of single-precision rounding.  Positions: Richardson-extrapolated central
    differences at steps s and 2s (s = 2^12 position units per coordinate step); masses: the loss is
    linear in them.  Bound: 1e-4 relative, no absolute floor."""
    import oracle_bind as ob
    from tigergrav_amd.solver import Solver

    gen, ewald, theta, leaf = FP64_CASES[case]
    x = gen()
    n = x.shape[0]
    h, m = ics.soft_len(n), 1.0 / n
    print(f"adjoint vs FP64 frozen tree {case}: N={n} ewald={ewald} theta={theta} leaf={leaf}", flush=True)
    s = Solver(ewald=ewald, soft_len=h, m=m, theta=theta, parts_per_node=leaf, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    rng = np.random.default_rng(len(case) + 200)
    w = rng.standard_normal((n, 3)).astype(np.float32)
    grad = s.adjoint(w)
    T = s.tree_arrays()
    order = T["order"]
    P = {"ewald": int(ewald), "soft_len": h, "m": m}
    wt = w[order].astype(np.float64)

    def loss64(pos_tree, mass_tree=None):
        return float(np.sum(wt * ob.frozen_solve64(P, T, pos_tree, mass_tree)[:, 1:]))

    d = rng.integers(-1, 2, size=x.shape).astype(np.int64)
    base = T["pos"].astype(np.int64)
    fd = []
    for st in (2 ** 12, 2 ** 13):
        lp = loss64((base + st * d[order]).astype(np.int32))
        lm = loss64((base - st * d[order]).astype(np.int32))
        fd.append((lp - lm) / (2 * st))
    fd_r = (4 * fd[0] - fd[1]) / 3  # per position unit
    pred = float(np.sum(grad[:, :3] * d)) * 2.0 ** -32
    rel = abs(fd_r - pred) / abs(pred)
    print(f"  positions: fd64 {fd_r:.9e} (steps s, 2s: {fd[0]:.9e}, {fd[1]:.9e}) adjoint {pred:.9e} rel {rel:.2e}",
          flush=True)
    assert rel <= 1e-4
    dm = m * rng.uniform(-1e-2, 1e-2, n)
    lp = loss64(T["pos"], (m + dm)[order])
    lm = loss64(T["pos"], (m - dm)[order])
    fdm = 0.5 * (lp - lm)
    predm = float(np.sum(grad[:, 3] * dm))
    relm = abs(fdm - predm) / abs(predm)
    print(f"  masses: fd64 {fdm:.9e} adjoint {predm:.9e} rel {relm:.2e}", flush=True)
    assert relm <= 1e-4
