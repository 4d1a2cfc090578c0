"""The plain drop-in: the reference's own solve code calling the GPU kernels through its own
interface.

oracle/_ref/libtgref_hip.so is oracle/ref_driver.cpp -- the reference's L2L (expansion::operator<<),
L2P (translate_L2) and kick_fmm call order -- with its eight gravity_* calls resolved by the
reference-signature shim tigergrav_amd/integration/gravity_hip.cpp, which forwards every call to
the C ABI (include/tg_gravity.h) and so to the HIP kernels, one synchronous call per node and
leaf from many OpenMP threads at once.  It is checked against the oracle solve (tgo_solve) on
identical lists, which is bit-identical to the same driver linked with the reference's own
src/gravity.cpp (tests/test_tree.py); the compiled reference itself is not used in-process here
because its kernels freeze their options at the first call (src/gravity.cpp:176-187), which would
carry one case's parameters into the next.  Tolerance: the kernels' parity bar, rms relative 1e-5
for g and for open-boundary phi; periodic phi bounded in absolute terms (see test_gpu_solve.py).

Both libraries are built in the container that holds the reference sources (make -C oracle ref)
and travel as built files; the test is skipped where they are absent.
"""
import math
import os
import time

import numpy as np
import pytest

import oracle_bind as ob

def _device_count():
    from tigergrav_amd._lib import lib

    return lib().tg_device_count()


pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(ob.REF_HIP_LIB),
                                 reason="oracle/_ref/libtgref_hip.so not built (needs the reference sources)")]


@pytest.fixture(autouse=True)
def _needs_device():
    # the shim aborts on a failed kernel call, as the reference's error paths do
    if _device_count() < 1:
        pytest.skip("no HIP device")

CASES = {
    "uniform_periodic": (lambda ics: ics.uniform(4096, 41), True, 0.7, 32),
    "plummer_open": (lambda ics: ics.plummer_spheres(8192, 42), False, 0.7, 64),
    "zeldovich_periodic": (lambda ics: ics.zeldovich(16, 43), True, 0.5, 64),
}


def rms_rel(a, b):
    return math.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2))


@pytest.mark.parametrize("case", list(CASES))
def test_reference_solve_on_gpu_kernels(case):
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    gen, ewald, theta, ppn = CASES[case]
    x = gen(ics)
    N = x.shape[0]
    h = ics.soft_len(N)
    print(f"drop-in {case}: N={N} ewald={ewald} theta={theta} leaf={ppn}", flush=True)
    s = Solver(ewald=ewald, soft_len=h, m=1.0 / N, theta=theta, parts_per_node=ppn, device=-1)
    s.set_particles(x)
    T = s.tree_arrays()
    gpu = ob.ref_lib(ob.REF_HIP_LIB)
    gpu.tgref_set_options(int(ewald), h, 1.0, N, theta)
    t0 = time.perf_counter()
    f_gpu, flop_gpu = ob.ref_solve(gpu, T)
    t1 = time.perf_counter()
    f_cpu, flop_cpu = ob.oracle_solve({"ewald": int(ewald), "soft_len": h, "m": 1.0 / N}, T)
    t2 = time.perf_counter()
    eg = rms_rel(f_gpu[:, 1:], f_cpu[:, 1:])
    dphi = math.sqrt(np.mean((f_gpu[:, 0] - f_cpu[:, 0]) ** 2))
    ep = rms_rel(f_gpu[:, 0], f_cpu[:, 0])
    print(f"  rms rel g {eg:.2e} phi {ep:.2e} (|dphi| {dphi:.2e}); drop-in {t1 - t0:.3f} s, oracle {t2 - t1:.3f} s",
          flush=True)
    assert flop_gpu == flop_cpu  # the shim returns the reference's own flop count
    assert eg < 1e-5
    assert (dphi < 2e-6) if ewald else (ep < 1e-5)


def test_batched_beats_plain_dropin_2m():
    """The plain drop-in on the full 2M open Plummer workload (8 spheres of scale 0.02 plus a 20%
    uniform background, theta 0.7, leaf 64): the reference's driver calling the GPU kernels one
    node or leaf at a time, against the batched solve on the same particles.  Same forces (the
    batched solve keeps the lists and kernels, and changes only the summation order), and the
    batched form far faster per solve."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.plummer_spheres(2 ** 21, seed=2)
    N = x.shape[0]
    h = ics.soft_len(N)
    host = Solver(ewald=False, soft_len=h, m=1.0 / N, theta=0.7, parts_per_node=64, device=-1)
    host.set_particles(x)
    T = host.tree_arrays()
    gpu = ob.ref_lib(ob.REF_HIP_LIB)
    gpu.tgref_set_options(0, h, 1.0, N, 0.7)
    print(f"drop-in 2M open Plummer: N={N}", flush=True)
    t0 = time.perf_counter()
    f_drop, _ = ob.ref_solve(gpu, T)
    t_drop = time.perf_counter() - t0
    s = Solver(ewald=False, soft_len=h, m=1.0 / N, theta=0.7, parts_per_node=64, device=0)
    s.set_particles(x)
    s.solve()
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        s.solve()
    s.synchronize()
    t_batched = (time.perf_counter() - t0) / 5
    f = s.forces()[T["order"]]  # tree order, as the drop-in returns it
    eg, ep = rms_rel(f[:, 1:], f_drop[:, 1:]), rms_rel(f[:, 0], f_drop[:, 0])
    print(f"  drop-in {t_drop:.3f} s per solve, batched {t_batched * 1e3:.3f} ms ({t_drop / t_batched:.0f}x); "
          f"rms rel g {eg:.2e} phi {ep:.2e}", flush=True)
    assert eg < 1e-5 and ep < 1e-5
    assert t_batched < t_drop
