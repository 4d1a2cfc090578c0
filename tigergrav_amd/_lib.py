"""Loader for the in-tree HIP library (tigergrav_amd/lib/libtigergrav_amd.so).

The product path has no CPU fallback: if the library is missing or cannot be loaded this module
raises, and every entry point that needs the GPU reports the HIP error it got.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libtigergrav_amd.so")

MULTI_DTYPE = np.dtype([("m", np.float32, (17,)), ("x", np.int32, (3,))])
FORCE_DTYPE = np.dtype([("phi", np.float64), ("g", np.float64, (3,))])

TG_OK, TG_EINVAL, TG_EHIP, TG_ENODEV, TG_ENOMEM = 0, -1, -2, -3, -4


class TGError(RuntimeError):
    pass


class Params(C.Structure):
    """tg_params (include/tg_gravity.h)."""
    _fields_ = [("ewald", C.c_int), ("soft_len", C.c_double), ("m", C.c_double), ("device", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TGError(f"{LIB_PATH} is not built; run `python -m tigergrav_amd.build` (or __graft_entry__.build())")
    # One HIP runtime per process.  torch ships its own libamdhip64 (soname libamdhip64.so.7, but
    # its libraries ask for it as libamdhip64.so): loaded after this library's /opt/rocm runtime it
    # would be a second runtime, which cannot initialise the GPU the first one holds.  Loaded
    # first, its runtime is the one this library's libamdhip64.so.7 resolves to.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i64, ci, u64p = C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_uint64)
    P = C.POINTER(Params)
    L.tg_gravity_PP_direct.argtypes = [P, vp, vp, i64, vp, i64, ci, u64p]
    for n in ("PC_direct", "PC_ewald"):
        getattr(L, "tg_gravity_" + n).argtypes = [P, vp, vp, i64, vp, i64, u64p]
    L.tg_gravity_PP_ewald.argtypes = [P, vp, vp, i64, vp, i64, u64p]
    for n in ("CC_direct", "CC_ewald", "CP_direct", "CP_ewald"):
        getattr(L, "tg_gravity_" + n).argtypes = [P, vp, vp, vp, i64, u64p]
    for n in ("PP_direct", "PC_direct", "PC_ewald", "PP_ewald", "CC_direct", "CC_ewald", "CP_direct", "CP_ewald"):
        getattr(L, "tg_gravity_" + n).restype = ci
    L.tg_last_error.restype = C.c_char_p
    L.tg_device_count.restype = ci
    L.tg_version.restype = ci
    L.tg_ewald_near_separation.argtypes = [vp]
    L.tg_ewald_near_separation.restype = C.c_double
    L.tg_ewald_far_separation.argtypes = [vp, C.c_double]
    L.tg_ewald_far_separation.restype = C.c_double
    _lib = L
    return L


def check(rc, what=""):
    if rc != TG_OK:
        msg = lib().tg_last_error().decode(errors="replace")
        raise TGError(f"{what} failed with status {rc}: {msg}")


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)
