"""The time step's checker (tgo_kick / tgo_drift, oracle/tg_oracle.c: part_vect_kick and
part_vect_drift, src/part_vect.cpp:341-436,493-540) and the product's time-grid functions
(tigergrav/time.hpp), pinned by what the rules fix rather than by restating them:

* the time grid: min_rung(t) = 31 - trailing zeros of t, rung steps t_max / 2^r, inc by the finest
  active step, a rung's particles kicked on exactly its own sub-grid;
* a new rung brackets the criterion: dt_r <= min(dt_max, eta sqrt(h / 2.8 / |a|)) < 2 dt_r;
* the two half kicks of an active particle add up to g G (dt_old + dt_new) / 2; inactive ones stay;
* the statistics are the plain sums they name;
* drift moves by v dt and wraps into the unit box;
* leapfrog on a harmonic well conserves energy to second order in the step.
"""
import numpy as np
import pytest

import oracle_bind as ob
from tigergrav_amd import step


def test_time_grid():
    for t in (1, 2, 3, 4, 6, 8, 12, 1 << 20, (1 << 31), 0x80000001, 0xFFFFFFFE):
        tz = (t & -t).bit_length() - 1
        assert step.min_rung(t) == 31 - tz, t
    assert step.min_rung(0) == 0
    assert step.time_inc(0, 3) == 1 << 28 and step.time_inc(5, 31) == 6
    assert step.time_to_double(2.0, 1 << 30) == 1.0
    for r in range(31):
        assert step.rung_to_dt(1.0, r) == 2.0 ** -r
    rng = np.random.default_rng(1)
    for dt in rng.uniform(1e-8, 2.0, 200):
        r = step.dt_to_rung(1.0, dt)
        assert 2.0 ** -r <= dt and (r == 0 or 2.0 ** -(r - 1) > dt)
    assert step.dt_to_rung(1.0, 1e-12) == -1  # past rung 30
    # a run whose finest rung is 3: every time is a multiple of 2^28 and min_rung picks the rungs
    # whose sub-grid the time lies on
    t = 0
    seen = {r: 0 for r in range(4)}
    for _ in range(8):
        t = step.time_inc(t, 3)
        for r in range(4):
            if step.min_rung(t) <= r:
                seen[r] += 1
    assert seen == {0: 1, 1: 2, 2: 4, 3: 8}


def _params(n, **kw):
    return step.params(n, soft_len=1e-3, **kw)


def test_new_rungs_bracket_the_criterion():
    n = 500
    P = _params(n, t_max=1.0, dt_max=1.0 / 64)
    rng = np.random.default_rng(2)
    f = np.zeros((n, 4))
    f[:, 1:] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-2, 6, (n, 1))
    v = np.zeros((n, 3), np.float32)
    rung = np.zeros(n, np.int8)
    rc, st = ob.kick(P, f, v, rung, 0, True)
    assert rc == 0
    a = np.linalg.norm(f[:, 1:], axis=1)
    crit = np.minimum(P.dt_max, P.eta * np.sqrt(P.soft_len / 2.8 / a))
    r = rung.astype(int)
    dt = 2.0 ** -r
    assert np.all(r >= 1)
    free = r > 1  # rung 1 may be the floor of max(., 1)
    assert np.all(dt[free] <= crit[free]) and np.all(crit[free] < 2 * dt[free])
    assert st["max_rung"] == max(int(np.max(r)), 0) or st["max_rung"] == int(np.max(np.maximum(r, 0)))


def test_half_kicks_and_inactive_particles():
    n = 400
    P = _params(n, t_max=1.0, dt_max=1.0 / 16)
    rng = np.random.default_rng(3)
    f = np.zeros((n, 4))
    f[:, 1:] = rng.standard_normal((n, 3)) * 50.0
    v0 = rng.standard_normal((n, 3)).astype(np.float32)
    rung0 = rng.integers(0, 6, n).astype(np.int8)
    min_rung = 3
    v, rung = v0.copy(), rung0.copy()
    rc, st = ob.kick(P, f, v, rung, min_rung, False)
    assert rc == 0
    act = rung0 >= min_rung
    # inactive: untouched
    assert np.array_equal(v[~act], v0[~act]) and np.array_equal(rung[~act], rung0[~act])
    # active: closing half kick of the old rung, opening half kick of the new criterion rung (at
    # least min_rung); the stored rung is at most one below the old and at least 1
    a = np.linalg.norm(f[:, 1:], axis=1)
    crit = np.minimum(P.dt_max, P.eta * np.sqrt(P.soft_len / 2.8 / a))
    rn = np.array([max(step.dt_to_rung(1.0, c), min_rung) for c in crit])
    dv = f[:, 1:] * (0.5 * 2.0 ** -rung0.astype(float) + 0.5 * 2.0 ** -rn)[:, None]
    np.testing.assert_allclose(v[act], (v0 + dv)[act], rtol=0, atol=4e-6 * np.max(np.abs(v0 + dv)))
    assert np.array_equal(rung[act], np.maximum(np.maximum(rn, rung0.astype(int) - 1), 1)[act])
    assert st["max_rung"] == int(np.max(rn[act]))


def test_statistics_are_the_sums_they_name():
    n = 1000
    P = _params(n, G=2.0)
    rng = np.random.default_rng(4)
    f = rng.standard_normal((n, 4))
    v0 = rng.standard_normal((n, 3)).astype(np.float32)
    v, rung = v0.copy(), np.full(n, 2, np.int8)
    rc, st = ob.kick(P, f, v, rung, 0, True)
    v64 = v0.astype(np.float64)
    np.testing.assert_allclose(st["p"], v64.sum(0) / n, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(st["kin"], 0.5 / n * np.sum(v64 ** 2), rtol=1e-6)
    np.testing.assert_allclose(st["g"], 2.0 * f[:, 1:].sum(0) / n, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st["pot"], 0.5 / n * 2.0 * f[:, 0].sum(), rtol=1e-12)


def test_drift_moves_and_wraps():
    from kernel_cases import double_to_pos

    n = 300
    P = _params(n, t_max=1.0)
    rng = np.random.default_rng(5)
    xd = rng.uniform(0, 1, (n, 3))
    x = np.asarray(double_to_pos(xd), np.int32).reshape(n, 3)
    v = (rng.standard_normal((n, 3)) * 20).astype(np.float32)
    x1 = x.copy()
    ek = ob.drift(P, x1, v, 2)
    want = np.mod(x.astype(np.float64) * 2.0 ** -32 + 0.5 + v.astype(np.float64) * 0.25, 1.0)
    got = x1.astype(np.float64) * 2.0 ** -32 + 0.5
    d = np.abs(got - want)
    d = np.minimum(d, 1.0 - d)
    assert np.max(d) < 1e-7  # float v * dt and the 2^-32 grid
    np.testing.assert_allclose(ek, 0.5 / n * np.sum(v.astype(np.float64) ** 2), rtol=1e-6)


def test_leapfrog_energy_on_a_harmonic_well():
    """Kick-drift-kick with the reference's rung logic (fixed rung: dt_max binds) on g = -w^2 (x - c):
    the energy error over a period is second order in the step (it drops ~4x when the step halves)."""
    from kernel_cases import double_to_pos

    w = 2 * np.pi * 4.0
    errs = []
    for dt_max in (1.0 / 256, 1.0 / 512):
        n = 64
        P = _params(n, t_max=1.0, dt_max=dt_max, eta=1e6)
        rng = np.random.default_rng(6)
        c = 0.5
        xd = c + rng.uniform(-0.02, 0.02, (n, 3))
        x = np.asarray(double_to_pos(xd), np.int32).reshape(n, 3)
        v = np.zeros((n, 3), np.float32)
        rung = np.zeros(n, np.int8)

        def force(x):
            f = np.zeros((n, 4))
            f[:, 1:] = -w * w * (x.astype(np.float64) * 2.0 ** -32 + 0.5 - c)
            return f

        def energy(x, v):
            return np.sum(0.5 * v.astype(np.float64) ** 2 + 0.5 * w * w * (x.astype(np.float64) * 2.0 ** -32 + 0.5 - c) ** 2)

        e0 = energy(x, v)
        rc, st = ob.kick(P, force(x), v, rung, 0, False)
        t, mr = 0, st["max_rung"]
        emax = 0.0
        for _ in range(int(0.25 / dt_max)):  # a period of 1/4
            ob.drift(P, x, v, mr)
            t = step.time_inc(t, mr)
            rc, st = ob.kick(P, force(x), v, rung, step.min_rung(t), False)
            mr = st["max_rung"]
            # after a kick the velocity leads by the opening half kick; the energy takes the
            # synchronised one
            vs = v - (0.5 * dt_max * force(x)[:, 1:]).astype(np.float32)
            emax = max(emax, abs(energy(x, vs) - e0) / e0)
        errs.append(emax)
    assert errs[0] < 5e-3  # 2.4e-3 measured
    assert 3.5 < errs[0] / errs[1] < 4.5, errs
