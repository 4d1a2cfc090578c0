"""ctypes bindings for the oracle (oracle/liboracle.so) and, where built, the compiled reference
kernels (oracle/_ref/libtgref.so).  Test infrastructure: imported by tests/, the smoke check and
bench.py's cpu_baseline leg only.

Array conventions (match include/tg_gravity.h):
  positions : int32 [n, 3]  (fixed point, x = p * 2^-32 + 0.5)
  forces    : float64 [n, 4] = (phi, gx, gy, gz)  -- the reference `force` struct
  multipoles: structured [n] of (m: float32[17], x: int32[3])  -- the reference `multi_src`
  expansion : float64 [35]
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")

MULTI_DTYPE = np.dtype([("m", np.float32, (17,)), ("x", np.int32, (3,))])
assert MULTI_DTYPE.itemsize == 80


class Params(C.Structure):
    _fields_ = [("ewald", C.c_int), ("soft_len", C.c_double), ("m", C.c_double)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            build_oracle()
        L = C.CDLL(path)
        vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
        for name in ("PP_direct",):
            for pre in ("tgo_", "tgo64_"):
                f = getattr(L, pre + name)
                f.argtypes = [vp, vp, vp, i64, vp, i64, ci]
        for name in ("PC_direct", "PC_ewald", "PP_ewald"):
            for pre in ("tgo_", "tgo64_"):
                f = getattr(L, pre + name)
                f.argtypes = [vp, vp, vp, i64, vp, i64]
        for name in ("CC_direct", "CC_ewald", "CP_direct", "CP_ewald"):
            for pre in ("tgo_", "tgo64_"):
                f = getattr(L, pre + name)
                f.argtypes = [vp, vp, vp, vp, i64]
        for name in ("tgo_PP_direct", "tgo_PC_direct", "tgo_PC_ewald", "tgo_PP_ewald", "tgo_CC_direct", "tgo_CC_ewald",
                     "tgo_CP_direct", "tgo_CP_ewald"):
            getattr(L, name).restype = C.c_uint64
        for name in ("tgo64_PP_direct", "tgo64_PC_direct", "tgo64_PC_ewald", "tgo64_PP_ewald", "tgo64_CC_direct",
                     "tgo64_CC_ewald", "tgo64_CP_direct", "tgo64_CP_ewald"):
            getattr(L, name).restype = None
        L.tgo_green_direct.argtypes = [vp, vp]
        L.tgo_green_ewald.argtypes = [vp, vp]
        L.tgo64_green_direct.argtypes = [vp, vp]
        L.tgo64_green_ewald.argtypes = [vp, vp]
        L.tgo_erfcexp.argtypes = [C.c_float, vp]
        L.tgo_erfcexp.restype = C.c_float
        L.tgo_exp.argtypes = [C.c_float]
        L.tgo_exp.restype = C.c_float
        L.tgo_sin.argtypes = [C.c_float]
        L.tgo_sin.restype = C.c_float
        L.tgo_direct64.argtypes = [ci, C.c_double, vp, i64, vp, vp, i64, vp, vp]
        L.tgo_ewald_corr64.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


KINDS = ("PP_direct", "PC_direct", "CC_direct", "CP_direct", "PP_ewald", "PC_ewald", "CC_ewald", "CP_ewald")


def call_kernel(prefix, kind, params, x, y, f=None, L=None, do_phi=1, libobj=None):
    """Run one kernel of the oracle family `prefix` ("tgo_" or "tgo64_").
    PP/PC kinds: x int32[ns,3]; returns forces float64[ns,4] (accumulated onto f if given).
    CC/CP kinds: x int32[3];    returns expansion float64[35] (accumulated onto L if given)."""
    Lb = libobj or lib()
    P = Params(int(params["ewald"]), float(params["soft_len"]), float(params["m"]))
    fn = getattr(Lb, prefix + kind)
    if kind.startswith("PP") or kind.startswith("PC"):
        x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
        ns = x.shape[0]
        out = np.zeros((ns, 4), np.float64) if f is None else np.array(f, np.float64, copy=True).reshape(ns, 4)
        if kind.startswith("PP"):
            y = np.ascontiguousarray(y, np.int32).reshape(-1, 3)
            args = [C.byref(P), _p(out), _p(x), ns, _p(y), y.shape[0]]
            if kind == "PP_direct":
                args.append(int(do_phi))
        else:
            y = np.ascontiguousarray(y, MULTI_DTYPE)
            args = [C.byref(P), _p(out), _p(x), ns, _p(y), y.shape[0]]
        rc = fn(*args)
        return out, rc
    x = np.ascontiguousarray(x, np.int32).reshape(3)
    out = np.zeros(35, np.float64) if L is None else np.array(L, np.float64, copy=True).reshape(35)
    if kind.startswith("CC"):
        y = np.ascontiguousarray(y, MULTI_DTYPE)
    else:
        y = np.ascontiguousarray(y, np.int32).reshape(-1, 3)
    rc = fn(C.byref(P), _p(out), _p(x), _p(y), y.shape[0])
    return out, rc


def green_direct(dx, fp64=False):
    if fp64:
        d = np.ascontiguousarray(dx, np.float64)
        D = np.zeros(35, np.float64)
        lib().tgo64_green_direct(_p(d), _p(D))
    else:
        d = np.ascontiguousarray(dx, np.float32)
        D = np.zeros(35, np.float32)
        lib().tgo_green_direct(_p(d), _p(D))
    return D


def green_ewald(dx, fp64=False):
    if fp64:
        d = np.ascontiguousarray(dx, np.float64)
        D = np.zeros(35, np.float64)
        lib().tgo64_green_ewald(_p(d), _p(D))
    else:
        d = np.ascontiguousarray(dx, np.float32)
        D = np.zeros(35, np.float32)
        lib().tgo_green_ewald(_p(d), _p(D))
    return D


def direct64(periodic, h, xs, ys, m):
    xs = np.ascontiguousarray(xs, np.int32).reshape(-1, 3)
    ys = np.ascontiguousarray(ys, np.int32).reshape(-1, 3)
    m = np.ascontiguousarray(np.broadcast_to(m, (ys.shape[0],)), np.float64)
    phi = np.zeros(xs.shape[0], np.float64)
    g = np.zeros((xs.shape[0], 3), np.float64)
    lib().tgo_direct64(int(periodic), float(h), _p(xs), xs.shape[0], _p(ys), _p(m), ys.shape[0], _p(phi), _p(g))
    return phi, g


def ewald_corr64(dx):
    d = np.ascontiguousarray(dx, np.float64)
    phi = np.zeros(1, np.float64)
    g = np.zeros(3, np.float64)
    lib().tgo_ewald_corr64(_p(d), _p(phi), _p(g))
    return phi[0], g


# ---- compiled reference (this container only) ----
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libtgref.so")


# the same reference driver with its gravity_* calls bound to the GPU kernels through the
# reference-signature shim (tigergrav_amd/integration/gravity_hip.cpp): the plain drop-in
REF_HIP_LIB = os.path.join(ORACLE_DIR, "_ref", "libtgref_hip.so")


def ref_available():
    return os.path.exists(REF_LIB)


def ref_lib(path=REF_LIB):
    L = C.CDLL(path)
    vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
    L.tgref_set_options.argtypes = [ci, C.c_double, C.c_double, C.c_uint64, C.c_double]
    L.tgref_PP_direct.argtypes = [vp, vp, i64, vp, i64, ci]
    for n in ("PC_direct", "PC_ewald", "PP_ewald"):
        getattr(L, "tgref_" + n).argtypes = [vp, vp, i64, vp, i64]
    for n in ("CC_direct", "CC_ewald", "CP_direct", "CP_ewald"):
        getattr(L, "tgref_" + n).argtypes = [vp, vp, vp, i64]
    for n in KINDS:
        getattr(L, "tgref_" + n).restype = C.c_uint64
    return L


def call_ref(L, kind, x, y, f=None, Lx=None, do_phi=1):
    fn = getattr(L, "tgref_" + kind)
    if kind.startswith("PP") or kind.startswith("PC"):
        x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
        ns = x.shape[0]
        out = np.zeros((ns, 4), np.float64) if f is None else np.array(f, np.float64, copy=True).reshape(ns, 4)
        if kind.startswith("PP"):
            y = np.ascontiguousarray(y, np.int32).reshape(-1, 3)
            args = [_p(out), _p(x), ns, _p(y), y.shape[0]]
            if kind == "PP_direct":
                args.append(int(do_phi))
        else:
            y = np.ascontiguousarray(y, MULTI_DTYPE)
            args = [_p(out), _p(x), ns, _p(y), y.shape[0]]
        rc = fn(*args)
        return out, rc
    x = np.ascontiguousarray(x, np.int32).reshape(3)
    out = np.zeros(35, np.float64) if Lx is None else np.array(Lx, np.float64, copy=True).reshape(35)
    if kind.startswith("CC"):
        y = np.ascontiguousarray(y, MULTI_DTYPE)
    else:
        y = np.ascontiguousarray(y, np.int32).reshape(-1, 3)
    rc = fn(_p(out), _p(x), _p(y), y.shape[0])
    return out, rc


# ---- whole-solve restatements over exported tree arrays ----
LIST_NAMES = ("cc_d", "cp_d", "cc_e", "cp_e", "pc_d", "pp_d", "pc_e", "pp_e")


def _solve_args(T):
    keep = []

    def a(x, dt):
        x = np.ascontiguousarray(x, dt)
        keep.append(x)
        return x

    pos = a(T["pos"], np.int32)
    nr = a(T["node_range"], np.int32)
    nc = a(T["node_child"], np.int32)
    npar = a(T["node_parent"], np.int32)
    nx = a(T["node_x"], np.int32)
    nm = a(T["node_m"], np.float32)
    ptrs = (C.c_void_p * 8)(*[_p(a(T[n + "_ptr"], np.int64)) for n in LIST_NAMES])
    lsts = (C.c_void_p * 8)(*[_p(a(T[n], np.int32)) if len(T[n]) else _p(a(np.zeros(1), np.int32)) for n in LIST_NAMES])
    keep += [ptrs, lsts]
    return pos, nr, nc, npar, nx, nm, ptrs, lsts, keep


def oracle_solve(params, T, stride=1):
    """The reference's solve sequence with the oracle's kernels (tgo_solve); forces in tree order."""
    L = lib()
    L.tgo_solve.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64] + [C.c_void_p] * 5 + [C.c_void_p, C.c_void_p,
                                                                                                 C.c_int, C.c_void_p,
                                                                                                 C.c_void_p]
    L.tgo_solve.restype = C.c_uint64
    pos, nr, nc, npar, nx, nm, ptrs, lsts, keep = _solve_args(T)
    P = Params(int(params["ewald"]), float(params["soft_len"]), float(params["m"]))
    out = np.zeros((pos.shape[0], 4), np.float64)
    flop = L.tgo_solve(C.byref(P), pos.shape[0], _p(pos), nr.shape[0], _p(nr), _p(nc), _p(npar), _p(nx), _p(nm), ptrs,
                       lsts, int(stride), _p(out), None)
    return out, flop


def oracle_solve_leaves(params, T, leaves):
    """The oracle solve (tgo_solve_want) for the particles of `leaves` only: those leaves and every
    ancestor are evaluated, nothing else.  Returns (forces in tree order, mask of evaluated particles)."""
    L = lib()
    L.tgo_solve_want.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64] + [C.c_void_p] * 5 + [
        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tgo_solve_want.restype = C.c_uint64
    pos, nr, nc, npar, nx, nm, ptrs, lsts, keep = _solve_args(T)
    want = np.zeros(nr.shape[0], np.uint8)
    parent = np.asarray(T["node_parent"], np.int64)
    a = np.unique(np.asarray(leaves, np.int64))
    while len(a):
        want[a] = 1
        a = parent[a]
        a = np.unique(a[a >= 0])
    P = Params(int(params["ewald"]), float(params["soft_len"]), float(params["m"]))
    out = np.zeros((pos.shape[0], 4), np.float64)
    L.tgo_solve_want(C.byref(P), pos.shape[0], _p(pos), nr.shape[0], _p(nr), _p(nc), _p(npar), _p(nx), _p(nm), ptrs,
                     lsts, _p(want), _p(out))
    mask = np.zeros(pos.shape[0], bool)
    for lf in np.unique(np.asarray(leaves, np.int64)):
        b, e = nr[lf]
        mask[b:e] = True
    return out, mask


def frozen_solve64(params, T, pos_tree, mass_tree=None):
    """The FP64 frozen-tree solve (tgo64_frozen_solve) of the tree and lists in T at positions
    pos_tree (and masses mass_tree, None = params["m"] for all; tree order): moments recomputed in
    FP64 about the kept centres, tgo64_ kernels, FP64 L2L/L2P.  Forces in tree order."""
    L = lib()
    L.tgo64_frozen_solve.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64] + [C.c_void_p] * 4 + [
        C.c_void_p, C.c_void_p, C.c_void_p]
    L.tgo64_frozen_solve.restype = None
    T = dict(T)
    T["pos"] = pos_tree
    pos, nr, nc, npar, nx, nm, ptrs, lsts, keep = _solve_args(T)
    P = Params(int(params["ewald"]), float(params["soft_len"]), float(params["m"]))
    out = np.zeros((pos.shape[0], 4), np.float64)
    mp = None
    if mass_tree is not None:
        mass_tree = np.ascontiguousarray(mass_tree, np.float64)
        mp = _p(mass_tree)
    L.tgo64_frozen_solve(C.byref(P), pos.shape[0], _p(pos), mp, nr.shape[0], _p(nr), _p(nc), _p(npar), _p(nx), ptrs,
                         lsts, _p(out))
    return out


def ref_solve(R, T, stride=1):
    """The same solve with the compiled reference's own L2L/L2P and kernels (tgref_solve)."""
    R.tgref_solve.argtypes = [C.c_int64, C.c_void_p, C.c_int64] + [C.c_void_p] * 5 + [C.c_void_p, C.c_void_p, C.c_int,
                                                                                       C.c_void_p]
    R.tgref_solve.restype = C.c_uint64
    pos, nr, nc, npar, nx, nm, ptrs, lsts, keep = _solve_args(T)
    out = np.zeros((pos.shape[0], 4), np.float64)
    flop = R.tgref_solve(pos.shape[0], _p(pos), nr.shape[0], _p(nr), _p(nc), _p(npar), _p(nx), _p(nm), ptrs, lsts,
                         int(stride), _p(out))
    return out, flop


class StepParams(C.Structure):
    """tgo_step_params (oracle/tg_oracle.h)."""
    _fields_ = [("G", C.c_double), ("eta", C.c_double), ("soft_len", C.c_double), ("dt_max", C.c_double),
                ("t_max", C.c_double), ("m", C.c_double), ("problem_size", C.c_double), ("glass", C.c_int),
                ("cosmic", C.c_int), ("ewald", C.c_int), ("scale", C.c_double), ("kick_dt1", C.c_double * 31),
                ("kick_dt2", C.c_double * 31)]


def step_params(P):
    """The oracle's parameter block from the product's (tigergrav_amd.step.StepParams)."""
    Q = StepParams()
    for name, _ in StepParams._fields_:
        if name.startswith("kick_dt"):
            for r in range(31):
                getattr(Q, name)[r] = getattr(P, name)[r]
        else:
            setattr(Q, name, getattr(P, name))
    return Q


def kick(P, f, v, rung, min_rung, do_out):
    """tgo_kick on host arrays (v float32 [n, 3] and rung int8 [n] updated in place); returns
    (status, stats dict)."""
    L = lib()
    L.tgo_kick.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    L.tgo_kick.restype = C.c_int
    Q = step_params(P)
    f = np.ascontiguousarray(f, np.float64)
    st = np.zeros(9)
    rc = L.tgo_kick(C.byref(Q), v.shape[0], _p(f), _p(v), _p(rung), int(min_rung), int(do_out), _p(st))
    return rc, {"g": st[0:3].copy(), "p": st[3:6].copy(), "pot": st[6], "kin": st[7], "max_rung": int(st[8])}


def drift(P, x, v, max_rung, drift_dt=0.0, a1=1.0):
    """tgo_drift on host arrays (x int32 [n, 3] updated in place); returns the kinetic energy."""
    L = lib()
    L.tgo_drift.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double]
    L.tgo_drift.restype = C.c_double
    Q = step_params(P)
    v = np.ascontiguousarray(v, np.float32)
    return L.tgo_drift(C.byref(Q), x.shape[0], _p(x), _p(v), int(max_rung), float(drift_dt), float(a1))


def fof(pos_tree, L2):
    """tgo_fof: the reference's friends-of-friends relabelling to convergence (tree-order ids)."""
    L = lib()
    L.tgo_fof.argtypes = [C.c_int64, C.c_void_p, C.c_uint64, C.c_void_p]
    L.tgo_fof.restype = None
    pos = np.ascontiguousarray(pos_tree, np.int32)
    g = np.zeros(pos.shape[0], np.uint64)
    L.tgo_fof(pos.shape[0], _p(pos), int(L2), _p(g))
    return g


def fof_link2(n, link_len):
    """The reference's integer squared linking length (src/part_vect.cpp:86-88)."""
    L = int((n ** (-1.0 / 3.0) * link_len) * 4294967295.0)
    return L * L
