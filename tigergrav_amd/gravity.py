"""The eight interaction kernels under the reference's names (tigergrav/gravity.hpp:39-46).

Each function has the reference's argument meaning and accumulate semantics:
  gravity_PP_direct(f, x, y, do_phi)      f += PP(x <- y)        (src/gravity.cpp:171-300)
  gravity_PC_direct(f, x, y)              f += M2P(x <- y)        (src/gravity.cpp:302-366)
  gravity_CC_direct(L, x, y)              L += M2L(x <- y)        (src/gravity.cpp:368-426)
  gravity_CP_direct(L, x, y)              L += P2L(x <- y)        (src/gravity.cpp:428-486)
  gravity_PP_ewald / PC_ewald / CC_ewald / CP_ewald: the periodic corrections (src/gravity.cpp:491-785)
and returns the reference's weighted flop count.  Arrays:
  f : float64 [n, 4] (phi, gx, gy, gz) or a FORCE_DTYPE structured array -- updated in place
  x : int32 [n, 3] sink positions (PP/PC) or int32 [3] expansion centre (CC/CP)
  y : int32 [n, 3] source particles (PP/CP) or MULTI_DTYPE [n] multipoles (PC/CC)
  L : float64 [35] local expansion -- updated in place
The kernel parameters (ewald, soft_len, m) are given per call through `Options` rather than the
reference's process-wide options singleton.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import MULTI_DTYPE, Params, check, lib, ptr


@dataclass
class Options:
    ewald: bool = True
    soft_len: float = 0.02 * 4096 ** (-1.0 / 3.0)
    m: float = 1.0 / 4096
    device: int = 0

    def c(self):
        return Params(int(bool(self.ewald)), float(self.soft_len), float(self.m), int(self.device))


def _forces(f):
    if f.dtype != np.float64 or not f.flags.c_contiguous or f.shape[-1:] != (4,):
        if f.dtype.names:  # FORCE_DTYPE
            return f
        raise ValueError("f must be a C-contiguous float64 [n, 4] array")
    return f


def _pos(a, n3=True):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.reshape(-1, 3) if n3 else a.reshape(3)


def _multi(y):
    y = np.ascontiguousarray(y)
    if y.dtype != MULTI_DTYPE:
        y = y.astype(MULTI_DTYPE)
    return y


def _leaf(name, opts, f, x, y, extra=()):
    f = _forces(f)
    x = _pos(x)
    ns = x.shape[0]
    if f.shape[0] != ns:
        raise ValueError("f and x sizes differ")
    flop = C.c_uint64(0)
    P = opts.c()
    fn = getattr(lib(), "tg_gravity_" + name)
    rc = fn(C.byref(P), ptr(f), ptr(x), ns, ptr(y), y.shape[0], *extra, C.byref(flop))
    check(rc, name)
    return flop.value


def gravity_PP_direct(f, x, y, do_phi=True, opts=None):
    return _leaf("PP_direct", opts or Options(), f, x, _pos(y), (int(bool(do_phi)),))


def gravity_PC_direct(f, x, y, opts=None):
    return _leaf("PC_direct", opts or Options(), f, x, _multi(y))


def gravity_PP_ewald(f, x, y, opts=None):
    return _leaf("PP_ewald", opts or Options(), f, x, _pos(y))


def gravity_PC_ewald(f, x, y, opts=None):
    return _leaf("PC_ewald", opts or Options(), f, x, _multi(y))


def _node(name, opts, L, x, y):
    if L.dtype != np.float64 or L.shape != (35,) or not L.flags.c_contiguous:
        raise ValueError("L must be a contiguous float64 [35] array")
    x = _pos(x, n3=False)
    flop = C.c_uint64(0)
    P = opts.c()
    rc = getattr(lib(), "tg_gravity_" + name)(C.byref(P), ptr(L), ptr(x), ptr(y), y.shape[0], C.byref(flop))
    check(rc, name)
    return flop.value


def gravity_CC_direct(L, x, y, opts=None):
    return _node("CC_direct", opts or Options(), L, x, _multi(y))


def gravity_CP_direct(L, x, y, opts=None):
    return _node("CP_direct", opts or Options(), L, x, _pos(y))


def gravity_CC_ewald(L, x, y, opts=None):
    return _node("CC_ewald", opts or Options(), L, x, _multi(y))


def gravity_CP_ewald(L, x, y, opts=None):
    return _node("CP_ewald", opts or Options(), L, x, _pos(y))


KERNELS = {
    "PP_direct": gravity_PP_direct, "PC_direct": gravity_PC_direct, "CC_direct": gravity_CC_direct,
    "CP_direct": gravity_CP_direct, "PP_ewald": gravity_PP_ewald, "PC_ewald": gravity_PC_ewald,
    "CC_ewald": gravity_CC_ewald, "CP_ewald": gravity_CP_ewald,
}


def ewald_near_separation(x):
    return lib().tg_ewald_near_separation(ptr(np.ascontiguousarray(x, np.float64)))


def ewald_far_separation(x, r=0.0):
    return lib().tg_ewald_far_separation(ptr(np.ascontiguousarray(x, np.float64)), float(r))
