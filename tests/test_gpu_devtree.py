"""The tree and its multipoles built on the device (device_build_tree, solver knob "gpu_tree")
against the host build (Tree::build + Tree::compute_multipoles, tg_tree.cpp), which restates
src/tree.cpp:122-296 and src/part_vect.cpp:542-635,895-968: every array bit for bit -- particle
order, positions, masses, node ranges, children, parents, levels, box ids, centres, moments, radii,
particle boxes and counts -- and the forces of a solve over each tree."""
import numpy as np
import pytest

from tigergrav_amd import ics

pytestmark = pytest.mark.gpu

ARRAYS = ("pos", "order", "mass", "leaves", "node_range", "node_child", "node_parent", "node_level", "node_boxid",
          "node_x", "node_m", "node_r", "node_box", "node_n")


def _coincident(n_extra=3000):
    # 100 coincident particles: a leaf at the depth limit
    return np.concatenate([np.repeat(ics.uniform(1, 71), 100, axis=0), ics.uniform(n_extra, 72)])


CASES = {
    # name: (generator, ewald, theta, leaf)
    "uniform32k_leaf64": (lambda: ics.uniform(32768, 1), True, 0.7, 64),
    "zeldovich64_leaf16": (lambda: ics.zeldovich(64, seed=4), True, 0.7, 16),
    "coincident_leaf32": (_coincident, True, 0.7, 32),
    "plummer2M_open_leaf64": (lambda: ics.plummer_spheres(2 ** 21, seed=2), False, 0.7, 64),
    "zeldovich128_leaf32": (lambda: ics.zeldovich(128, seed=3, n_s=-2.0, rms_cells=0.3), True, 0.3, 32),
    "one_particle": (lambda: ics.uniform(1, 5), True, 0.7, 32),
}


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8) if a.dtype.kind == "f" else a


def _tree(x, ewald, theta, leaf, gpu_tree):
    from tigergrav_amd.solver import Solver

    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=ics.soft_len(n), m=1.0 / n, theta=theta, parts_per_node=leaf, device=0)
    s.tune("gpu_tree", gpu_tree)
    s.set_particles(x)
    return s


@pytest.mark.parametrize("case", list(CASES))
def test_device_tree_is_bit_identical(case):
    gen, ewald, theta, leaf = CASES[case]
    x = gen()
    print(f"device tree {case}: N={x.shape[0]} leaf={leaf}", flush=True)
    sh = _tree(x, ewald, theta, leaf, 0)
    sd = _tree(x, ewald, theta, leaf, 1)
    ih, idv = sh.info(), sd.info()
    print(f"  nodes {ih['n_nodes']} leaves {ih['n_leaves']} levels {ih['max_level']}; host tree {ih['t_tree']:.4f} s "
          f"multipoles {ih['t_multipoles']:.4f} s, device {idv['t_tree']:.4f} s + {idv['t_multipoles']:.4f} s", flush=True)
    for name in ARRAYS:
        a, b = sh.export(name), sd.export(name)
        assert a.shape == b.shape, name
        assert np.array_equal(_bits(a), _bits(b)), name
    assert ih["count"] == idv["count"] and ih["n_nodes"] == idv["n_nodes"] and ih["max_level"] == idv["max_level"]
    if x.shape[0] <= 40000:
        for s in (sh, sd):
            s.solve()
            s.synchronize()
        assert np.array_equal(_bits(sh.forces()), _bits(sd.forces()))
    sh.close()
    sd.close()


def test_device_tree_16m_sample():
    """16,777,216 Zel'dovich particles (P(k) ~ k^-1.5, rms 0.5 cell), leaf 32: the whole particle order
    and every node array bit for bit."""
    x = ics.zeldovich(256, seed=5, n_s=-1.5, rms_cells=0.5)
    print(f"device tree zeldovich256: N={x.shape[0]}", flush=True)
    sh = _tree(x, True, 0.7, 32, 0)
    hs = {k: sh.export(k) for k in ARRAYS}
    ih = sh.info()
    sh.close()
    sd = _tree(x, True, 0.7, 32, 1)
    idv = sd.info()
    print(f"  nodes {ih['n_nodes']}; host {ih['t_tree']:.3f} + {ih['t_multipoles']:.3f} s, device {idv['t_tree']:.3f} + "
          f"{idv['t_multipoles']:.3f} s", flush=True)
    for name in ARRAYS:
        b = sd.export(name)
        assert hs[name].shape == b.shape and np.array_equal(_bits(hs[name]), _bits(b)), name
    sd.close()
