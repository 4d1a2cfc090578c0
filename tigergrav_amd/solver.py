"""Batched device-resident force solve (include/tg_solver.h) from Python.

    s = Solver(ewald=True, soft_len=h, m=1/N, theta=0.7, parts_per_node=64)
    s.set_particles(x)        # host: tree, multipoles, interaction lists; device: upload
    s.solve(); s.synchronize()
    f = s.forces()            # float64 [N, 4] = (phi, gx, gy, gz), input order

device=-1 builds the tree and lists on the host only (export works, solve does not).
"""
import ctypes as C

import numpy as np

from ._lib import Params, TGError, lib, ptr

KINDS = ("PP_direct", "PC_direct", "CC_direct", "CP_direct", "PP_ewald", "PC_ewald", "CC_ewald", "CP_ewald")
LIST_NAMES = ("cc_d", "cp_d", "cc_e", "cp_e", "pc_d", "pp_d", "pc_e", "pp_e")


class SolveInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("n_nodes", C.c_int64), ("n_leaves", C.c_int64), ("max_level", C.c_int64),
                ("count", C.c_uint64 * 8), ("items", C.c_int64 * 8), ("list_len", C.c_int64 * 8),
                ("list_bytes", C.c_int64), ("t_tree", C.c_double), ("t_multipoles", C.c_double),
                ("t_walk", C.c_double), ("t_upload", C.c_double), ("t_detail", C.c_double * 6)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
    L.tg_solver_create.argtypes = [C.POINTER(Params), C.c_double, ci, C.POINTER(vp)]
    L.tg_solver_destroy.argtypes = [vp]
    L.tg_solver_set_particles.argtypes = [vp, vp, vp, i64]
    L.tg_solver_solve.argtypes = [vp]
    L.tg_solver_solve_graph.argtypes = [vp]
    L.tg_solver_update_particles.argtypes = [vp, vp, vp]
    L.tg_solver_adjoint.argtypes = [vp, vp, vp, vp]
    L.tg_solver_tune.argtypes = [vp, C.c_char_p, ci]
    L.tg_solver_synchronize.argtypes = [vp]
    L.tg_solver_get_forces.argtypes = [vp, vp]
    L.tg_solver_kernel_ms.argtypes = [vp, vp]
    L.tg_solver_collect_ms.argtypes = [vp, vp, vp]
    L.tg_solver_info.argtypes = [vp, C.POINTER(SolveInfo)]
    L.tg_solver_export.argtypes = [vp, C.c_char_p, vp, i64]
    L.tg_solver_export.restype = i64
    L.tg_solver_last_error.restype = C.c_char_p
    for n in ("create", "destroy", "set_particles", "solve", "solve_graph", "tune", "update_particles", "adjoint", "synchronize", "get_forces", "kernel_ms", "collect_ms", "info"):
        getattr(L, "tg_solver_" + n).restype = ci
    _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().tg_solver_last_error().decode(errors="replace")
        raise TGError(f"{what} failed with status {rc}: {msg}")


class Solver:
    def __init__(self, ewald=True, soft_len=1e-3, m=1e-3, theta=0.7, parts_per_node=64, device=0):
        L = _bind()
        self.params = Params(int(bool(ewald)), float(soft_len), float(m), int(device))
        self.theta = float(theta)
        self.parts_per_node = int(parts_per_node)
        h = C.c_void_p()
        _check(L.tg_solver_create(C.byref(self.params), self.theta, self.parts_per_node, C.byref(h)), "tg_solver_create")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().tg_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, x, mass=None):
        x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
        self.n = x.shape[0]
        mp = None
        if mass is not None:
            mass = np.ascontiguousarray(mass, np.float32).reshape(-1)
            mp = ptr(mass)
        _check(lib().tg_solver_set_particles(self._h, ptr(x), mp, self.n), "tg_solver_set_particles")

    def update_particles(self, x, mass=None):
        """New positions (and masses) for the same particles, input order, evaluated on the kept
        tree, lists and expansion centres (moments recomputed about those centres)."""
        x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
        if x.shape[0] != self.n:
            raise ValueError("update_particles needs the same particle count")
        mp = None
        if mass is not None:
            mass = np.ascontiguousarray(mass, np.float32).reshape(-1)
            mp = ptr(mass)
        _check(lib().tg_solver_update_particles(self._h, ptr(x), mp), "tg_solver_update_particles")

    def adjoint(self, w, with_time=False):
        """Gradients of L = sum_i w_i . g_i for the current particle set with its tree, lists and
        expansion centres held fixed: (N, 4) array of (dL/dx, dL/dy, dL/dz, dL/dm), input order,
        positions in box units.  With with_time, also the adjoint's device time in ms."""
        w = np.ascontiguousarray(w, np.float32).reshape(-1, 3)
        if w.shape[0] != self.n:
            raise ValueError("w needs one row per particle")
        g = np.zeros((self.n, 4), np.float64)
        ms = np.zeros(1, np.float64)
        _check(lib().tg_solver_adjoint(self._h, ptr(w), ptr(g), ptr(ms)), "tg_solver_adjoint")
        return (g, float(ms[0])) if with_time else g

    def find_groups(self, link_len):
        """Friends-of-friends groups of the current particles (tree::find_groups, src/tree.cpp:637-760):
        (group id per particle in input order, number of groups); link_len in units of the mean
        particle spacing (the reference's b).  Ids are tree-order indices, as the reference's."""
        L = lib()
        L.tg_solver_find_groups.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_int64)]
        g = np.zeros(self.n, np.uint64)
        ng = C.c_int64()
        _check(L.tg_solver_find_groups(self._h, float(link_len), g.ctypes.data, C.byref(ng)), "tg_solver_find_groups")
        return g, ng.value

    def solve(self):
        _check(lib().tg_solver_solve(self._h), "tg_solver_solve")

    def tune(self, knob, value):
        """Set a tuning knob: "pp_persistent" (0/1), "node_chunk" (1..4096, from the next
        set_particles), "resident" (work groups of the persistent PP kernel)."""
        _check(lib().tg_solver_tune(self._h, knob.encode(), int(value)), "tg_solver_tune")

    def solve_graph(self):
        """The same solve replayed from a captured HIP graph (no per-kind timings)."""
        _check(lib().tg_solver_solve_graph(self._h), "tg_solver_solve_graph")

    def synchronize(self):
        _check(lib().tg_solver_synchronize(self._h), "tg_solver_synchronize")

    def forces(self):
        f = np.zeros((self.n, 4), np.float64)
        _check(lib().tg_solver_get_forces(self._h, ptr(f)), "tg_solver_get_forces")
        return f

    def kernel_ms(self):
        ms = np.zeros(9, np.float64)
        _check(lib().tg_solver_kernel_ms(self._h, ptr(ms)), "tg_solver_kernel_ms")
        out = {k: float(ms[i]) for i, k in enumerate(KINDS)}
        out["local"] = float(ms[8])
        return out

    def collect_ms(self):
        """Mean device time (ms) per solve of each kind's kernels over every solve enqueued since
        the previous call, and the number of those solves."""
        ms = np.zeros(9, np.float64)
        n = np.zeros(1, np.int64)
        _check(lib().tg_solver_collect_ms(self._h, ptr(ms), ptr(n)), "tg_solver_collect_ms")
        k = max(int(n[0]), 1)
        out = {name: float(ms[i]) / k for i, name in enumerate(KINDS)}
        out["local"] = float(ms[8]) / k
        return out, int(n[0])

    def info(self):
        s = SolveInfo()
        _check(lib().tg_solver_info(self._h, C.byref(s)), "tg_solver_info")
        return {
            "n": s.n, "n_nodes": s.n_nodes, "n_leaves": s.n_leaves, "max_level": s.max_level,
            "count": {k: int(s.count[i]) for i, k in enumerate(KINDS)},
            "items": {k: int(s.items[i]) for i, k in enumerate(KINDS)},
            "list_len": {k: int(s.list_len[i]) for i, k in enumerate(KINDS)},
            "list_bytes": s.list_bytes, "t_tree": s.t_tree, "t_multipoles": s.t_multipoles, "t_walk": s.t_walk,
            "t_upload": s.t_upload,
            "t_detail": dict(zip(("walk_setup", "walk_steps", "walk_assembly", "census", "prepare", "transfer"),
                                 (float(v) for v in s.t_detail))),
        }

    _DT = {"pos": (np.int32, 3), "order": (np.int64, 1), "mass": (np.float32, 1), "leaves": (np.int32, 1),
           "node_range": (np.int32, 2), "node_child": (np.int32, 2), "node_parent": (np.int32, 1),
           "node_level": (np.int32, 1), "node_x": (np.int32, 3), "node_m": (np.float32, 17), "node_r": (np.float32, 1), "node_box": (np.float64, 6), "node_n": (np.uint64, 1),
           "node_boxid": (np.uint64, 2), "pp_d_items": (np.int32, 4), "pp_d_list": (np.int32, 1)}

    def export(self, name):
        L = lib()
        nb = L.tg_solver_export(self._h, name.encode(), None, 0)
        if nb < 0:
            raise TGError(f"unknown export {name}")
        if name in self._DT:
            dt, w = self._DT[name]
        elif name.endswith("_ptr"):
            dt, w = np.int64, 1
        else:
            dt, w = np.int32, 1
        a = np.zeros(nb // np.dtype(dt).itemsize, dt)
        L.tg_solver_export(self._h, name.encode(), ptr(a), nb)
        return a.reshape(-1, w) if w > 1 else a

    def tree_arrays(self):
        """Everything a CPU re-evaluation of the solve needs (tree order)."""
        d = {k: self.export(k) for k in ("pos", "order", "node_range", "node_child", "node_parent", "node_x", "node_m",
                                         "node_r", "node_level", "leaves")}
        for n in LIST_NAMES:
            d[n] = self.export(n)
            d[n + "_ptr"] = self.export(n + "_ptr")
        return d
