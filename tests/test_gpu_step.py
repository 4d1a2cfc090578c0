"""The time step on the device (tg_kick, tg_drift; include/tg_step.h) against the oracle's
restatement of part_vect_kick / part_vect_drift (src/part_vect.cpp:341-436,493-540): velocities,
rungs and positions bit for bit; the statistics (summed in another order) to 1e-12 relative.  And
the device-resident step loop (tigergrav_amd.step.Stepper: tree, solve, kick and drift with the
particle state in device memory) against the same loop driven from the host through the solver's
host interface and the oracle's kick and drift: identical trajectories."""
import numpy as np
import pytest

import oracle_bind as ob
from tigergrav_amd import ics, step

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.as_tensor(a).to("cuda:0")


def _close(a, b, rel=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rel * np.maximum(np.abs(b), 1e-300) + 1e-300)


@pytest.mark.parametrize("mode", ["newton", "newton_ewald", "glass", "cosmic"])
def test_kick_drift_match_oracle(mode):
    n = 20000
    rng = np.random.default_rng({"newton": 1, "newton_ewald": 2, "glass": 3, "cosmic": 4}[mode])
    kw = {}
    if mode == "newton_ewald":
        kw = {"ewald": True, "kick_dt2": rng.uniform(0, 1e-3, 31)}
    elif mode == "glass":
        kw = {"glass": True}
    elif mode == "cosmic":
        kw = {"cosmic": True, "ewald": True, "scale": 0.37, "kick_dt1": rng.uniform(0, 1e-3, 31),
              "kick_dt2": rng.uniform(0, 1e-3, 31)}
    P = step.params(n, soft_len=ics.soft_len(n), G=1.3, t_max=1.0, **kw)
    x = ics.uniform(n, 11)
    f = np.zeros((n, 4))
    f[:, 0] = rng.standard_normal(n)
    f[:, 1:] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-1, 4, (n, 1))
    v = rng.standard_normal((n, 3)).astype(np.float32)
    rung = rng.integers(0, 8, n).astype(np.int8)
    print(f"kick/drift {mode}: N={n}", flush=True)
    for min_rung, do_out in ((0, True), (4, False), (2, True)):
        vo, ro = v.copy(), rung.copy()
        rc, so = ob.kick(P, f, vo, ro, min_rung, do_out)
        assert rc == 0
        vd, rd, fd = _dev(v), _dev(rung), _dev(f)
        sd = step.kick(P, fd, vd, rd, min_rung, do_out)
        assert np.array_equal(vd.cpu().numpy().view(np.uint32), vo.view(np.uint32)), (min_rung, do_out)
        assert np.array_equal(rd.cpu().numpy(), ro)
        assert sd["max_rung"] == so["max_rung"]
        for k in ("g", "p", "pot", "kin"):
            assert _close(sd[k], so[k]), (k, sd[k], so[k])
        v, rung = vo, ro
    for mr in (0, 3, 9):
        xo = x.copy()
        eo = ob.drift(P, xo, v, mr, drift_dt=2.5e-3, a1=0.8)
        xd = _dev(x)
        ed = step.drift(P, xd, _dev(v), mr, drift_dt=2.5e-3, a1=0.8)
        assert np.array_equal(xd.cpu().numpy(), xo), mr
        assert _close(ed, eo)
        x = xo


def test_kick_rejects_a_rung_off_the_grid():
    n = 64
    P = step.params(n, soft_len=1e-3)
    f = np.zeros((n, 4))
    f[:, 1] = 1.0
    rung = np.zeros(n, np.int8)
    rung[5] = 40
    with pytest.raises(Exception, match="outside 0..30"):
        step.kick(P, _dev(f), _dev(np.zeros((n, 3), np.float32)), _dev(rung), 0, False)


@pytest.mark.parametrize("ewald", [False, True])
def test_device_resident_loop_matches_host_loop(ewald):
    """Six steps of the loop from a cold Plummer-like clump (open) or a uniform box (periodic): the
    device loop (positions, velocities, rungs and forces never leave device memory) and a host loop
    (host arrays through tg_solver_set_particles / get_forces, the oracle's kick and drift) visit the
    same states bit for bit."""
    from tigergrav_amd.solver import Solver

    n = 4096
    x0 = ics.plummer_spheres(n, 21) if not ewald else ics.uniform(n, 22)
    v0 = (np.random.default_rng(23).standard_normal((n, 3)) * 0.05).astype(np.float32)
    h = ics.soft_len(n)
    st = step.Stepper(x0, v0, ewald=ewald, theta=0.6, leaf=32, t_max=1.0, dt_max=1.0 / 64)
    print(f"stepper ewald={ewald}: N={n}", flush=True)
    # host loop
    s = Solver(ewald=ewald, soft_len=h, m=1.0 / n, theta=0.6, parts_per_node=32, device=0)
    P = st.P
    x, v, rung = x0.copy(), v0.copy(), np.zeros(n, np.int8)

    def host_solve(x):
        s.set_particles(x)
        s.solve()
        s.synchronize()
        return s.forces()

    rc, sh = ob.kick(P, host_solve(x), v, rung, 0, True)
    sd = st.first()
    assert sh["max_rung"] == sd["max_rung"]
    itime, mr = 0, sh["max_rung"]
    for k in range(6):
        ob.drift(P, x, v, mr)
        itime = step.time_inc(itime, mr)
        rc, sh = ob.kick(P, host_solve(x), v, rung, step.min_rung(itime), False)
        mr = sh["max_rung"]
        sd = st.step()
        xd, vd, rd = st.state()
        assert st.itime == itime
        assert np.array_equal(xd, x), k
        assert np.array_equal(vd.view(np.uint32), v.view(np.uint32)), k
        assert np.array_equal(rd, rung), k
        assert sd["max_rung"] == mr
    print(f"  after 6 steps: t = {st.t:.6f}, max rung {mr}, rungs used {sorted(set(rung.tolist()))}", flush=True)
