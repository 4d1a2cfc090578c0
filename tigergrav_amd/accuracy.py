"""Force accuracy of a solve: FP64 direct sums on the GPU and the reference's error statistics.

The reference's solver test (src/main.cpp:137-164) compares the tree solve with a direct sum over
all particles and prints error statistics (src/output.cpp:17-51).  `direct_forces` is that direct
sum (include/tg_direct.h: FP64, converged Ewald sum when periodic) for a sample of sinks;
`error_stats` reduces the per-particle relative acceleration errors |a - a_ref| / |a_ref|.
"""
import ctypes as C

import numpy as np

from ._lib import Params, TGError, lib, ptr

_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.tg_direct_forces.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                       C.c_void_p]
        L.tg_direct_forces.restype = C.c_int
        L.tg_direct_gradient.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_int64, C.c_void_p]
        L.tg_direct_gradient.restype = C.c_int
        L.tg_direct_last_error.restype = C.c_char_p
        _bound = True
    return L


def direct_forces(x, sinks, ewald, soft_len, m, mass=None, device=0):
    """(len(sinks), 4) array of (phi, gx, gy, gz) at particles `sinks` from all particles `x`
    (int32 fixed point, (N, 3)), FP64 on the GPU."""
    L = _bind()
    x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
    sinks = np.ascontiguousarray(sinks, np.int64).reshape(-1)
    out = np.zeros((sinks.size, 4), np.float64)
    mp = None
    if mass is not None:
        mass = np.ascontiguousarray(mass, np.float32).reshape(-1)
        mp = ptr(mass)
    prm = Params(int(bool(ewald)), float(soft_len), float(m), int(device))
    rc = L.tg_direct_forces(C.byref(prm), ptr(x), mp, x.shape[0], ptr(sinks), sinks.size, ptr(out))
    if rc != 0:
        raise TGError(f"tg_direct_forces failed with status {rc}: {L.tg_direct_last_error().decode(errors='replace')}")
    return out


def direct_gradient(x, w, sinks, ewald, soft_len, m, mass=None, device=0):
    """(len(sinks), 4) array of (dL/dx, dL/dy, dL/dz, dL/dm) at particles `sinks` for the direct-sum
    loss L = sum_i w_i . g_i over all particles (positions in box units), FP64 on the GPU."""
    L = _bind()
    x = np.ascontiguousarray(x, np.int32).reshape(-1, 3)
    w = np.ascontiguousarray(w, np.float32).reshape(-1, 3)
    sinks = np.ascontiguousarray(sinks, np.int64).reshape(-1)
    out = np.zeros((sinks.size, 4), np.float64)
    mp = None
    if mass is not None:
        mass = np.ascontiguousarray(mass, np.float32).reshape(-1)
        mp = ptr(mass)
    prm = Params(int(bool(ewald)), float(soft_len), float(m), int(device))
    rc = L.tg_direct_gradient(C.byref(prm), ptr(x), mp, x.shape[0], ptr(w), ptr(sinks), sinks.size, ptr(out))
    if rc != 0:
        raise TGError(f"tg_direct_gradient failed with status {rc}: {L.tg_direct_last_error().decode(errors='replace')}")
    return out


def error_stats(f, f_ref):
    """Relative acceleration errors of forces f against f_ref, both (n, 4) (phi, g): RMS, mean,
    99th percentile and maximum of |g - g_ref| / |g_ref|, and the RMS relative potential error."""
    f = np.asarray(f, np.float64)
    f_ref = np.asarray(f_ref, np.float64)
    da = np.linalg.norm(f[:, 1:4] - f_ref[:, 1:4], axis=1)
    a = np.linalg.norm(f_ref[:, 1:4], axis=1)
    rel = da / a
    dp = np.abs(f[:, 0] - f_ref[:, 0]) / np.abs(f_ref[:, 0])
    return {"rms": float(np.sqrt(np.mean(rel ** 2))), "mean": float(np.mean(rel)), "p99": float(np.percentile(rel, 99)),
            "max": float(np.max(rel)), "phi_rms": float(np.sqrt(np.mean(dp ** 2))), "n": int(rel.size)}
