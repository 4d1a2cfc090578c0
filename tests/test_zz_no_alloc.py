"""tg_gravity_PP_direct_noalloc -- the particle-particle drop-in that allocates nothing (its staging
lives in a fixed device array) -- against tg_gravity_PP_direct, the entry point it mirrors, through
the C ABI: the same bits on two calls in a row (the second smaller, so stale contents of the
shared array would show), periodic and open, with and without the potential, onto non-zero forces;
and inputs beyond the array's limits are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from tigergrav_amd import ics
from tigergrav_amd._lib import TG_EINVAL, Params, lib

MAX_SINKS, MAX_SOURCES = 16384, 262144


def _bind():
    L = lib()
    vp, i64, ci, u64p = C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_uint64)
    L.tg_gravity_PP_direct_noalloc.argtypes = [C.POINTER(Params), vp, vp, i64, vp, i64, ci, u64p]
    L.tg_gravity_PP_direct_noalloc.restype = ci
    return L


def _call(fn, P, f, x, y, do_phi):
    flop = C.c_uint64(0)
    rc = fn(C.byref(P), f.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), x.shape[0],
            y.ctypes.data_as(C.c_void_p), y.shape[0], int(do_phi), C.byref(flop))
    return rc, flop.value


@pytest.mark.gpu
@pytest.mark.parametrize("ewald", [True, False])
@pytest.mark.parametrize("do_phi", [True, False])
def test_noalloc_matches_pp_direct(ewald, do_phi):
    L = _bind()
    n = 4096
    P = Params(int(ewald), float(ics.soft_len(n)), 1.0 / n, 0)
    rng = np.random.default_rng(17 + 2 * int(ewald) + int(do_phi))
    # two calls in a row; the second with fewer sinks and sources
    for ns, nj in ((200, 5000), (17, 300)):
        x = ics.uniform(ns, int(rng.integers(1 << 30)))
        y = np.concatenate([x, ics.uniform(nj - ns, int(rng.integers(1 << 30)))])
        f0 = rng.standard_normal((ns, 4))
        fa, fb = f0.copy(), f0.copy()
        print(f"noalloc ewald={ewald} do_phi={do_phi}: ns={ns} nj={nj}", flush=True)
        rca, flopa = _call(L.tg_gravity_PP_direct, P, fa, x, y, do_phi)
        rcb, flopb = _call(L.tg_gravity_PP_direct_noalloc, P, fb, x, y, do_phi)
        assert rca == 0 and rcb == 0, (rca, rcb, L.tg_last_error())
        assert flopa == flopb
        np.testing.assert_array_equal(fb, fa)
        assert np.any(fb != f0)


@pytest.mark.gpu
def test_noalloc_no_sources_applies_phi_self():
    L = _bind()
    n = 4096
    P = Params(1, float(ics.soft_len(n)), 1.0 / n, 0)
    x = ics.uniform(5, 3)
    fa, fb = np.zeros((5, 4)), np.zeros((5, 4))
    y = np.zeros((0, 3), np.int32)
    assert _call(L.tg_gravity_PP_direct, P, fa, x, y, True)[0] == 0
    assert _call(L.tg_gravity_PP_direct_noalloc, P, fb, x, y, True)[0] == 0
    np.testing.assert_array_equal(fb, fa)


def test_noalloc_refuses_sizes_beyond_its_array():
    """Refused with TG_EINVAL and a message naming the limit, before any device call (no GPU needed)."""
    L = _bind()
    P = Params(1, 1e-3, 1e-3, 0)
    f = np.zeros((MAX_SINKS + 1, 4))
    x = np.zeros((MAX_SINKS + 1, 3), np.int32)
    rc, _ = _call(L.tg_gravity_PP_direct_noalloc, P, f, x, x[:1], True)
    assert rc == TG_EINVAL and b"TG_NOALLOC_MAX_SINKS" in L.tg_last_error()
    y = np.zeros((MAX_SOURCES + 1, 3), np.int32)
    rc, _ = _call(L.tg_gravity_PP_direct_noalloc, P, f[:1], x[:1], y, True)
    assert rc == TG_EINVAL and b"TG_NOALLOC_MAX_SOURCES" in L.tg_last_error()
