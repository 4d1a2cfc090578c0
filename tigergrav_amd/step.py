"""The time step on the device (include/tg_step.h): kick and drift with hierarchical rungs, and a
stepper that runs the reference's main loop (src/main.cpp:233-330) with the particle state kept in
device memory.

    st = Stepper(x, v, ewald=True, theta=0.5, leaf=32)   # x int32 [N, 3], v float32 [N, 3]
    st.run(t_end)                                        # solve, kick, drift, rebuild ... until t_max
    x, v, rung = st.state()                               # host copies

Arrays handed to kick / drift are torch tensors on the solver's device (torch is used for device
memory only); per-particle results are the reference's, operation for operation.
"""
import ctypes as C

import numpy as np

from ._lib import TGError, lib

RUNG_MAX = 30


class StepParams(C.Structure):
    """tg_step_params (include/tg_step.h): the options part_vect_kick / part_vect_drift read."""
    _fields_ = [("G", C.c_double), ("eta", C.c_double), ("soft_len", C.c_double), ("dt_max", C.c_double),
                ("t_max", C.c_double), ("m", C.c_double), ("problem_size", C.c_double), ("glass", C.c_int),
                ("cosmic", C.c_int), ("ewald", C.c_int), ("scale", C.c_double),
                ("kick_dt1", C.c_double * (RUNG_MAX + 1)), ("kick_dt2", C.c_double * (RUNG_MAX + 1)),
                ("device", C.c_int)]


class KickStats(C.Structure):
    _fields_ = [("g", C.c_double * 3), ("p", C.c_double * 3), ("pot", C.c_double), ("kin", C.c_double),
                ("max_rung", C.c_int32)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
    L.tg_kick.argtypes = [C.POINTER(StepParams), i64, vp, vp, vp, ci, ci, C.POINTER(KickStats)]
    L.tg_kick.restype = ci
    L.tg_drift.argtypes = [C.POINTER(StepParams), i64, vp, vp, ci, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.tg_drift.restype = ci
    L.tg_rung_to_dt.argtypes = [C.c_double, ci]
    L.tg_rung_to_dt.restype = C.c_double
    L.tg_dt_to_rung.argtypes = [C.c_double, C.c_double]
    L.tg_dt_to_rung.restype = ci
    L.tg_min_rung.argtypes = [C.c_uint32]
    L.tg_min_rung.restype = ci
    L.tg_time_inc.argtypes = [C.c_uint32, ci]
    L.tg_time_inc.restype = C.c_uint32
    L.tg_time_to_double.argtypes = [C.c_double, C.c_uint32]
    L.tg_time_to_double.restype = C.c_double
    L.tg_solver_set_particles_device.argtypes = [vp, vp, vp, i64]
    L.tg_solver_set_particles_device.restype = ci
    L.tg_solver_get_forces_device.argtypes = [vp, vp]
    L.tg_solver_get_forces_device.restype = ci
    _bound = True
    return L


def params(n, soft_len, ewald=False, G=1.0, eta=0.2, t_max=1.0, dt_max=None, m=None, glass=False, cosmic=False,
           scale=1.0, kick_dt1=None, kick_dt2=None, device=0):
    """tg_step_params with the reference's defaults (src/options.cpp: eta 0.2, dt_max = t_max / 64,
    m = 1 / N)."""
    P = StepParams()
    P.G, P.eta, P.soft_len = G, eta, soft_len
    P.t_max = t_max
    P.dt_max = t_max / 64.0 if dt_max is None else dt_max
    P.m = 1.0 / n if m is None else m
    P.problem_size = float(n)
    P.glass, P.cosmic, P.ewald = int(glass), int(cosmic), int(ewald)
    P.scale = scale
    for r in range(RUNG_MAX + 1):
        P.kick_dt1[r] = 0.0 if kick_dt1 is None else kick_dt1[r]
        P.kick_dt2[r] = 0.0 if kick_dt2 is None else kick_dt2[r]
    P.device = device
    return P


def _check(rc, what):
    if rc != 0:
        raise TGError(f"{what} failed with status {rc}: {lib().tg_last_error().decode(errors='replace')}")


def kick(P, f, v, rung, min_rung, do_out):
    """part_vect_kick on device tensors: f float64 [N, 4] (phi, g), v float32 [N, 3], rung int8 [N]
    (updated in place).  Returns the statistics as a dict."""
    L = _bind()
    n = v.shape[0]
    st = KickStats()
    _check(L.tg_kick(C.byref(P), n, f.data_ptr(), v.data_ptr(), rung.data_ptr(), int(min_rung), int(do_out),
                     C.byref(st)), "tg_kick")
    return {"g": np.array(st.g[:]), "p": np.array(st.p[:]), "pot": st.pot, "kin": st.kin, "max_rung": int(st.max_rung)}


def drift(P, x, v, max_rung, drift_dt=0.0, a1=1.0):
    """part_vect_drift on device tensors: x int32 [N, 3] (updated in place), v float32 [N, 3].
    Returns the kinetic energy before the move."""
    L = _bind()
    ek = C.c_double()
    _check(L.tg_drift(C.byref(P), x.shape[0], x.data_ptr(), v.data_ptr(), int(max_rung), float(drift_dt), float(a1),
                      C.byref(ek)), "tg_drift")
    return ek.value


def min_rung(t):
    return _bind().tg_min_rung(int(t) & 0xFFFFFFFF)


def time_inc(t, max_rung):
    return _bind().tg_time_inc(int(t) & 0xFFFFFFFF, int(max_rung))


def time_to_double(t_max, t):
    return _bind().tg_time_to_double(float(t_max), int(t) & 0xFFFFFFFF)


def rung_to_dt(t_max, rung):
    return _bind().tg_rung_to_dt(float(t_max), int(rung))


def dt_to_rung(t_max, dt):
    return _bind().tg_dt_to_rung(float(t_max), float(dt))


class Stepper:
    """The reference's Newtonian main loop (src/main.cpp:233-330 without cosmology, output or
    groups): a force solve on the tree of the current positions, a kick of the active rungs, a
    drift by the finest active step, a rebuild; the integer time advances by the finest step.
    Positions, velocities, rungs and forces stay in device memory throughout."""

    def __init__(self, x, v, ewald=True, theta=0.7, leaf=64, soft_len=None, G=1.0, eta=0.2, t_max=1.0, dt_max=None,
                 device=0, graph=False):
        import torch

        from . import ics
        from .solver import Solver

        _bind()
        n = x.shape[0]
        self.n = n
        self.dev = torch.device(f"cuda:{device}")
        self.x = torch.as_tensor(np.ascontiguousarray(x, np.int32).reshape(n, 3)).to(self.dev)
        self.v = torch.as_tensor(np.ascontiguousarray(v, np.float32).reshape(n, 3)).to(self.dev)
        self.rung = torch.zeros(n, dtype=torch.int8, device=self.dev)
        self.f = torch.zeros((n, 4), dtype=torch.float64, device=self.dev)
        h = ics.soft_len(n) if soft_len is None else soft_len
        self.P = params(n, h, ewald=ewald, G=G, eta=eta, t_max=t_max, dt_max=dt_max, device=device)
        self.solver = Solver(ewald=ewald, soft_len=h, m=1.0 / n, theta=theta, parts_per_node=leaf, device=device)
        self.solver.tune("gpu_tree", 1)
        self.graph = graph
        self.itime = 0
        self.t = 0.0
        self.max_rung = 0
        self.steps = 0

    def _solve(self):
        import torch

        torch.cuda.synchronize(self.dev)
        L = lib()
        h = self.solver._h
        _check(L.tg_solver_set_particles_device(h, self.x.data_ptr(), None, self.n), "tg_solver_set_particles_device")
        (self.solver.solve_graph if self.graph else self.solver.solve)()
        _check(L.tg_solver_get_forces_device(h, self.f.data_ptr()), "tg_solver_get_forces_device")

    def first(self):
        """The opening solve and kick at t = 0 (every particle active, src/main.cpp:254-266)."""
        self._solve()
        st = kick(self.P, self.f, self.v, self.rung, min_rung(0), True)
        self.max_rung = st["max_rung"]
        return st

    def step(self):
        """One pass of the loop body (src/main.cpp:266-318): drift by the finest active step,
        rebuild and solve, kick the rungs active at the new time."""
        ekin = drift(self.P, self.x, self.v, self.max_rung)
        self.itime = time_inc(self.itime, self.max_rung)
        self._solve()
        st = kick(self.P, self.f, self.v, self.rung, min_rung(self.itime), False)
        self.max_rung = st["max_rung"]
        self.t = time_to_double(self.P.t_max, self.itime)
        self.steps += 1
        st["ekin_before_drift"] = ekin
        return st

    def run(self, t_end):
        if self.steps == 0 and self.itime == 0:
            self.first()
        while self.t < t_end:
            self.step()

    def state(self):
        return self.x.cpu().numpy(), self.v.cpu().numpy(), self.rung.cpu().numpy()
