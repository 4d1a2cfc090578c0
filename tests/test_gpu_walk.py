"""The device walk (tg_solver_tune(s, "gpu_walk", 1), tigergrav_amd/csrc/tg_walk.hip) against the
host walk (Tree::walk, itself pinned to the reference's lists by tests/test_tree.py): every list
and every CSR pointer bit-identical, the same interaction census, and a solve over the device
lists bit-identical to one over the host lists."""
import numpy as np
import pytest

from tigergrav_amd import ics
from tigergrav_amd.solver import LIST_NAMES, Solver

pytestmark = pytest.mark.gpu

CASES = {
    # name: (generator, ewald, theta, leaf)
    "uniform_periodic": (lambda: ics.uniform(4096, 31), True, 0.7, 32),
    "plummer_open": (lambda: ics.plummer_spheres(8192, 32), False, 0.7, 64),
    "zeldovich_periodic_t05": (lambda: ics.zeldovich(32, 33), True, 0.5, 32),
    "plummer_periodic_t03": (lambda: ics.plummer_spheres(4096, 34), True, 0.3, 32),
    "uniform_open_t04_leaf8": (lambda: ics.uniform(3000, 35), False, 0.4, 8),
    "clustered_periodic_t07": (lambda: ics.plummer_spheres(65536, 36), True, 0.7, 64),
}


def _build(x, ewald, theta, leaf, device, gpu_walk):
    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=ics.soft_len(n), m=1.0 / n, theta=theta, parts_per_node=leaf, device=device)
    if device >= 0:
        s.tune("gpu_walk", 1 if gpu_walk else 0)
    s.set_particles(x)
    return s


@pytest.mark.parametrize("case", list(CASES))
def test_device_walk_matches_host_walk(case):
    gen, ewald, theta, leaf = CASES[case]
    x = gen()
    print(f"walk {case}: N={x.shape[0]} ewald={ewald} theta={theta} leaf={leaf}", flush=True)
    host = _build(x, ewald, theta, leaf, -1, False)
    dev = _build(x, ewald, theta, leaf, 0, True)
    for name in LIST_NAMES:
        for suffix in ("", "_ptr"):
            a, b = host.export(name + suffix), dev.export(name + suffix)
            assert a.shape == b.shape and np.array_equal(a, b), (name + suffix, a.shape, b.shape)
    assert host.info()["count"] == dev.info()["count"]
    assert sum(dev.info()["count"].values()) > 0


def test_solve_over_device_lists_is_bit_identical():
    x = ics.zeldovich(24, 37)
    a = _build(x, True, 0.6, 32, 0, False)
    b = _build(x, True, 0.6, 32, 0, True)
    for s in (a, b):
        s.solve()
        s.synchronize()
    assert np.array_equal(a.forces(), b.forces())


# The device walk's own lists at the benchmark's size, checked against the reference's opening
# criterion by the FP64 verifier (tests/mac_verify.py, which builds no lists) and for exact
# coverage of every sink: the 128^3 Zel'dovich set at leaf 32, theta 0.7 (every entry) and theta
# 0.3 (the entries of 3000 seeded sink leaves and their ancestors; coverage on 1000 of them).
@pytest.mark.parametrize("theta,sample", [(0.7, None), (0.3, 3000)])
def test_device_walk_lists_pass_mac_verifier_2m(theta, sample):
    from mac_verify import MacVerifier, coverage_exact

    x = ics.zeldovich(128, 3)
    n = x.shape[0]
    print(f"device walk 2M zeldovich theta={theta} leaf=32: verifying", flush=True)
    s = _build(x, True, theta, 32, 0, True)
    T = s.tree_arrays()
    for name in LIST_NAMES:
        T[name] = s.export(name)
        T[name + "_ptr"] = s.export(name + "_ptr")
    leaves = T["leaves"]
    rng = np.random.default_rng(5)
    sinks = None if sample is None else rng.choice(leaves, sample, replace=False)
    bad = MacVerifier(T, theta, ics.soft_len(n), True).check(sink_leaves=sinks)
    assert all(v == 0 for v in bad.values()), bad
    cov = rng.choice(leaves, 1000 if sample else 4000, replace=False)
    assert coverage_exact(T, cov, "d") == [] and coverage_exact(T, cov, "e") == []
    assert sum(len(T[k]) for k in LIST_NAMES) > 0
