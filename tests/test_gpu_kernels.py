"""GPU parity of the eight per-call kernels (C ABI) against the oracle.

Two references on identical inputs (tests/kernel_cases.py, both kernel-parameter sets):
  * tgo_   -- the reference's float kernels restated (bit-exact with the compiled reference);
  * tgo64_ -- the same formulas in FP64.
Metric: rms relative error over a case set, sqrt(sum |d|^2 / sum |ref|^2), taken separately for
accelerations and potentials (PP/PC) or over all 35 coefficients (CC/CP).  Tolerance: 1e-5.
For the Ewald kinds the comparison with tgo64_ takes the correction alone (its own norm as the
scale).  The reference's float evaluation (tgo_) forms the correction with a cancellation at close
range that costs it about 1e-5 of its own, so against tgo_ the scale is the total the correction is
added to (correction + the matching direct kernel's term).
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from kernel_cases import PARAMS, kernel_cases

pytestmark = pytest.mark.gpu

TOL = 1e-5
DIRECT_OF = {"PP_ewald": "PP_direct", "PC_ewald": "PC_direct", "CC_ewald": "CC_direct", "CP_ewald": "CP_direct"}


@pytest.fixture(scope="module")
def tg():
    from tigergrav_amd import gravity

    return gravity


def run_gpu(tg, kind, P, c):
    opts = tg.Options(ewald=bool(P["ewald"]), soft_len=P["soft_len"], m=P["m"])
    if kind[:2] in ("PP", "PC"):
        f = np.zeros((np.asarray(c["x"]).reshape(-1, 3).shape[0], 4))
        if kind == "PP_direct":
            tg.KERNELS[kind](f, c["x"], c["y"], c["do_phi"], opts=opts)
        else:
            tg.KERNELS[kind](f, c["x"], c["y"], opts=opts)
        return f
    L = np.zeros(35)
    tg.KERNELS[kind](L, c["x"], c["y"], opts=opts)
    return L


def run_oracle(prefix, kind, P, c):
    if kind[:2] in ("PP", "PC"):
        return ob.call_kernel(prefix, kind, P, c["x"], c["y"], do_phi=c["do_phi"])[0]
    return ob.call_kernel(prefix, kind, P, c["x"], c["y"])[0]


def rel(num, den):
    return math.sqrt(num / den) if den > 0 else math.sqrt(num)


@pytest.mark.parametrize("cfg", ["periodic", "open"])
@pytest.mark.parametrize("kind", ob.KINDS)
def test_kernel_parity(tg, cfg, kind):
    P = PARAMS[cfg]
    acc = {"g32": [0.0, 0.0], "g64": [0.0, 0.0], "p32": [0.0, 0.0], "p64": [0.0, 0.0]}
    for c in kernel_cases(kind, P):
        if c["name"].endswith(("symmetry_points", "coincident", "spline_edges")):
            continue  # exact-limit cases: test_kernel_limits
        gpu = run_gpu(tg, kind, P, c)
        r32 = run_oracle("tgo_", kind, P, c)
        r64 = run_oracle("tgo64_", kind, P, c)
        if kind[:2] in ("PP", "PC"):
            parts = {"g": (slice(1, 4),), "p": (slice(0, 1),)}
        else:
            parts = {"g": (slice(0, 35),), "p": None}
        extra = None
        if kind in DIRECT_OF:
            extra = run_oracle("tgo64_", DIRECT_OF[kind], dict(P, ewald=1), c)
        for key, sl in parts.items():
            if sl is None:
                continue
            if key == "p" and kind == "PP_direct" and not c["do_phi"]:
                continue
            a, b32, b64 = gpu[..., sl[0]], r32[..., sl[0]], r64[..., sl[0]]
            den64 = np.sum(b64 ** 2)
            den32 = den64 + (np.sum(extra[..., sl[0]] ** 2) if extra is not None else 0.0)
            acc[key + "32"][0] += np.sum((a - b32) ** 2)
            acc[key + "32"][1] += den32
            acc[key + "64"][0] += np.sum((a - b64) ** 2)
            acc[key + "64"][1] += den64
    errs = {key: rel(num, den) for key, (num, den) in acc.items() if den > 0}
    print(f"{kind} {cfg}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for key, e in errs.items():
        assert e < TOL, f"{kind} {cfg} {key}: rms rel {e:.3e}"


@pytest.mark.parametrize("kind", ["PC_ewald", "CC_ewald", "CP_ewald"])
def test_ewald_correction_alone(tg, kind):
    """Well-separated sources: the correction itself agrees with the FP64 formula."""
    P = PARAMS["periodic"]
    num = den = 0.0
    for c in kernel_cases(kind, P):
        if c["name"].endswith("coincident"):
            continue
        gpu = run_gpu(tg, kind, P, c)
        r64 = run_oracle("tgo64_", kind, P, c)
        sl = slice(1, 4) if kind[:2] == "PC" else slice(0, 35)
        num += np.sum((gpu[..., sl] - r64[..., sl]) ** 2)
        den += np.sum(r64[..., sl] ** 2)
    print(f"{kind} correction alone vs FP64: rms rel {rel(num, den):.2e}", flush=True)
    assert rel(num, den) < TOL


def test_pp_ewald_close_pairs_beat_reference_float(tg):
    """For close pairs the GPU correction (analytic n = 0 image) is closer to the FP64 formula than
    the reference's float evaluation is."""
    P = PARAMS["periodic"]
    e_gpu = e_ref = den = 0.0
    for c in kernel_cases("PP_ewald", P):
        if c["name"].endswith("symmetry_points"):
            continue
        gpu = run_gpu(tg, "PP_ewald", P, c)[:, 1:]
        r32 = run_oracle("tgo_", "PP_ewald", P, c)[:, 1:]
        r64 = run_oracle("tgo64_", "PP_ewald", P, c)[:, 1:]
        e_gpu += np.sum((gpu - r64) ** 2)
        e_ref += np.sum((r32 - r64) ** 2)
        den += np.sum(r64 ** 2)
    print(f"PP_ewald vs FP64: GPU {rel(e_gpu, den):.2e}, reference float {rel(e_ref, den):.2e}", flush=True)
    assert rel(e_gpu, den) < TOL
    assert e_gpu < e_ref


def test_kernel_limits(tg):
    """Spline switch points, symmetry points and coincident sources."""
    P = PARAMS["periodic"]
    for kind in ("PP_direct", "PP_ewald", "CP_ewald"):
        for c in kernel_cases(kind, P):
            if not c["name"].endswith(("symmetry_points", "coincident", "spline_edges")):
                continue
            gpu = run_gpu(tg, kind, P, c)
            r64 = run_oracle("tgo64_", kind, P, c)
            r32 = run_oracle("tgo_", kind, P, c)
            scale = max(np.max(np.abs(r64)), 1e-30)
            np.testing.assert_allclose(gpu, r64, rtol=2e-5, atol=2e-5 * scale, err_msg=c["name"])
            if kind == "PP_direct":
                np.testing.assert_allclose(gpu, r32, rtol=2e-5, atol=2e-5 * scale, err_msg=c["name"])


def test_accumulate_and_flops_match_reference(tg):
    """Outputs are added to the caller's arrays and the reference's flop counts are returned.
    Values are checked against the FP64 formula (tgo64_) plus the initial contents; the
    reference's own float Ewald output carries ~1e-5 errors of its own."""
    P = PARAMS["periodic"]
    opts = tg.Options(ewald=True, soft_len=P["soft_len"], m=P["m"])
    for kind in ob.KINDS:
        c = kernel_cases(kind, P)[1]
        if kind[:2] in ("PP", "PC"):
            f = np.array(c["f0"], np.float64, copy=True)
            args = (f, c["x"], c["y"]) + ((c["do_phi"],) if kind == "PP_direct" else ())
            flop = tg.KERNELS[kind](*args, opts=opts)
            rflop = ob.call_kernel("tgo_", kind, P, c["x"], c["y"], f=c["f0"], do_phi=c["do_phi"])[1]
            ref = ob.call_kernel("tgo64_", kind, P, c["x"], c["y"], f=c["f0"], do_phi=c["do_phi"])[0]
            contrib = ref - c["f0"]
        else:
            f = np.array(c["L0"], np.float64, copy=True)
            flop = tg.KERNELS[kind](f, c["x"], c["y"], opts=opts)
            rflop = ob.call_kernel("tgo_", kind, P, c["x"], c["y"], L=c["L0"])[1]
            ref = ob.call_kernel("tgo64_", kind, P, c["x"], c["y"], L=c["L0"])[0]
            contrib = ref - c["L0"]
        np.testing.assert_allclose(f, ref, rtol=1e-5, atol=2e-5 * np.max(np.abs(contrib)), err_msg=kind)
        assert flop == rflop, kind


def test_empty_inputs_and_phi_self(tg):
    P = PARAMS["periodic"]
    opts = tg.Options(ewald=True, soft_len=P["soft_len"], m=P["m"])
    x = np.array([[1, 2, 3], [4, 5, 6]], np.int32)
    f = np.zeros((2, 4))
    # no sources: PP_direct still applies -phi_self (src/gravity.cpp:290-296)
    assert tg.gravity_PP_direct(f, x, np.zeros((0, 3), np.int32), True, opts=opts) == 69 * 2
    np.testing.assert_allclose(f[:, 0], 315.0 / 128.0 * P["m"] / P["soft_len"])
    f2 = np.zeros((0, 4))
    assert tg.gravity_PC_direct(f2, np.zeros((0, 3), np.int32), np.zeros(3, tg.MULTI_DTYPE), opts=opts) == 0
    L = np.zeros(35)
    assert tg.gravity_CC_ewald(L, np.zeros(3, np.int32), np.zeros(0, ob.MULTI_DTYPE), opts=opts) == 0
    assert np.all(L == 0)


def test_invalid_arguments_are_errors(tg):
    import ctypes as C

    from tigergrav_amd._lib import TG_EINVAL, Params, lib

    P = Params(1, 1e-3, 1e-3, 0)
    flop = C.c_uint64(0)
    assert lib().tg_gravity_PP_direct(C.byref(P), None, None, -1, None, 0, 1, C.byref(flop)) == TG_EINVAL
    assert lib().tg_gravity_CC_direct(C.byref(P), None, None, None, 5, C.byref(flop)) == TG_EINVAL
    P.device = 999
    f = np.zeros((1, 4))
    x = np.zeros((1, 3), np.int32)
    rc = lib().tg_gravity_PP_direct(C.byref(P), f.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), 1,
                                    x.ctypes.data_as(C.c_void_p), 1, 1, C.byref(flop))
    assert rc < 0
