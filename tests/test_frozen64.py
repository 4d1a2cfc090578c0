"""The FP64 frozen-tree solve of the oracle (tgo64_frozen_solve), the reference point of the
adjoint's derivative checks, pinned on the CPU:

* at the build's own positions it equals the oracle's float solve of the same lists (the
  reference's solve sequence, tgo_solve) to float rounding;
* its moments follow the particles: moving every particle by a common offset moves nothing
  relative to the kept centres' frame but the centres' offsets, and the forces of a translated set
  on a tree whose centres are translated too are unchanged (open boundary, exact integer shift);
* it is smooth: central differences at two step sizes agree to second order (no rounding noise at
  the 1e-8 level), which is what makes it a derivative reference.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob


def build(x, ewald, theta, leaf):
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=ics.soft_len(n), m=1.0 / n, theta=theta, parts_per_node=leaf, device=-1)
    s.set_particles(x)
    return s.tree_arrays(), {"ewald": int(ewald), "soft_len": ics.soft_len(n), "m": 1.0 / n}


CASES = {
    "plummer_open": (lambda ics: ics.plummer_spheres(4096, 82), False, 0.7, 32),
    "zeldovich_periodic_leaf8": (lambda ics: ics.zeldovich(16, 83), True, 0.7, 8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_frozen64_equals_float_solve_at_build_positions(case):
    from tigergrav_amd import ics

    gen, ewald, theta, leaf = CASES[case]
    T, P = build(gen(ics), ewald, theta, leaf)
    f64 = ob.frozen_solve64(P, T, T["pos"])
    f32, _ = ob.oracle_solve(P, T)
    e = math.sqrt(np.sum((f64[:, 1:] - f32[:, 1:]) ** 2) / np.sum(f32[:, 1:] ** 2))
    assert e < 1e-5, e


def test_frozen64_is_smooth():
    """Directional central differences at steps 4s, 2s, s converge as s^2: each halving of the step
    quarters the change (ratio 4 within 10%), down to changes of 5e-7 of the derivative -- no
    rounding floor, so its central differences are a derivative reference to that level."""
    from tigergrav_amd import ics

    x = ics.plummer_spheres(2048, 84)
    T, P = build(x, False, 0.7, 32)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((x.shape[0], 3))
    d = rng.integers(-1, 2, size=T["pos"].shape).astype(np.int64)

    def loss(step):
        return float(np.sum(w * ob.frozen_solve64(P, T, (T["pos"] + step * d).astype(np.int32))[:, 1:]))

    fd = [(loss(s) - loss(-s)) / (2 * s) for s in (2 ** 12, 2 ** 11, 2 ** 10)]
    r = (fd[0] - fd[1]) / (fd[1] - fd[2])
    assert 3.6 < r < 4.4, r
    assert abs(fd[1] - fd[2]) < 1e-5 * abs(fd[2])
