"""The solve's work built on the device from the device tree and the device walk's lists
(tigergrav_amd/csrc/tg_prep.hip) against the work the host builds from the host walk's lists and
a host copy of the tree (prepare_work, tigergrav_amd/csrc/tg_upload.cpp): on the same particles
every work array -- leaf items in their sorted order, expanded PP and CP lists with PP direct's
near prefix first, the CC/CP waves, the partial slot of every entry, each slot's node, each node's
first slot -- and every tree array the solve reads (particle records, multipoles, centres, levels,
parents, leaves, root paths, each particle's leaf, the input order) is identical element for
element, the interaction census agrees, and the solves give the same forces bit for bit.  The two
walks give identical lists (tests/test_gpu_walk.py), so any difference is the preparation's."""
import numpy as np
import pytest

from tigergrav_amd import ics
from tigergrav_amd.solver import LIST_NAMES, Solver

pytestmark = pytest.mark.gpu

NODE_KINDS = (2, 3, 6, 7)  # CC/CP direct, CC/CP Ewald (tg_solve_info order)
LEAF_KINDS = (0, 1, 4, 5)  # PP/PC direct, PP/PC Ewald

CASES = {
    # name: (generator, ewald, theta, leaf, node_chunk)
    "uniform_periodic_t07": (lambda: ics.uniform(20000, 41), True, 0.7, 32, 64),
    "zeldovich_periodic_t03_leaf16": (lambda: ics.zeldovich(40, 42), True, 0.3, 16, 64),
    "plummer_open_t05_chunk17": (lambda: ics.plummer_spheres(30000, 43), False, 0.5, 64, 17),
    "plummer_periodic_t04_leaf8_chunk32": (lambda: ics.plummer_spheres(12000, 44), True, 0.4, 8, 32),
    "zeldovich_periodic_t03_chunk256": (lambda: ics.zeldovich(32, 47), True, 0.3, 16, 256),
}


def _build(x, ewald, theta, leaf, chunk, gpu_walk):
    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=ics.soft_len(n), m=1.0 / n, theta=theta, parts_per_node=leaf, device=0)
    s.tune("gpu_walk", 1 if gpu_walk else 0)
    s.tune("node_chunk", chunk)
    s.set_particles(x)
    return s


def _work(s):
    d = {}
    for k in NODE_KINDS + LEAF_KINDS:
        d[f"work_items_{k}"] = s.export(f"work_items_{k}")
        d[f"work_list_{k}"] = s.export(f"work_list_{k}")
        if k in NODE_KINDS:
            d[f"work_owner_{k}"] = s.export(f"work_owner_{k}")
    d["slot_node"] = s.export("slot_node")
    d["slot_ptr"] = s.export("slot_ptr")
    # the tree arrays the solve reads (made on the device from a device-kept tree, on the host
    # from a host tree)
    for name in ("parts", "multis", "nodepos", "level", "parent", "leaf_items", "path", "pleaf", "order"):
        d["solve_" + name] = s.export("solve_" + name)
    return d


@pytest.mark.parametrize("case", list(CASES))
def test_device_work_matches_host_work(case):
    gen, ewald, theta, leaf, chunk = CASES[case]
    x = gen()
    print(f"prep {case}: N={x.shape[0]} ewald={ewald} theta={theta} leaf={leaf} chunk={chunk}", flush=True)
    host = _build(x, ewald, theta, leaf, chunk, False)
    dev = _build(x, ewald, theta, leaf, chunk, True)
    a, b = _work(host), _work(dev)
    for name in a:
        assert a[name].shape == b[name].shape and np.array_equal(a[name], b[name]), (name, a[name].shape, b[name].shape)
    ia, ib = host.info(), dev.info()
    assert ia["count"] == ib["count"]
    assert ia["items"] == ib["items"] and ia["list_len"] == ib["list_len"]
    for key in ("n_nodes", "n_leaves", "max_level"):
        assert ia[key] == ib[key], key
    # PP direct carries a near prefix somewhere, and the waves pack more than one chunk somewhere
    pp = b["work_items_0"].reshape(-1, 4)
    assert pp.shape[0] > 0 and np.any((pp[:, 1] >> 7) > 0)
    for s in (host, dev):
        s.solve()
        s.synchronize()
    assert np.array_equal(host.forces(), dev.forces())
    # the host copy of the device walk's lists, made on demand, is the host walk's
    for name in LIST_NAMES:
        assert np.array_equal(host.export(name), dev.export(name)), name
    host.close()
    dev.close()


def test_rebuild_after_device_work_keeps_lists_and_forces():
    """A second set_particles on the same solver replaces the device lists and their views: the
    forces equal those of a fresh solver, and the exported lists are the new tree's."""
    x1 = ics.zeldovich(24, 45)
    x2 = ics.plummer_spheres(9000, 46)
    s = _build(x1, True, 0.5, 32, 64, True)
    s.solve()
    s.synchronize()
    s.export("pp_d")  # a host copy of the first tree's lists
    n2 = x2.shape[0]
    s.set_particles(x2)
    s.solve()
    s.synchronize()
    f = s.forces()
    t = Solver(ewald=True, soft_len=ics.soft_len(x1.shape[0]), m=1.0 / x1.shape[0], theta=0.5, parts_per_node=32, device=0)
    t.tune("node_chunk", 64)  # as s was built
    t.set_particles(x2)
    t.solve()
    t.synchronize()
    assert np.array_equal(f, t.forces())
    assert np.array_equal(s.export("pp_d"), t.export("pp_d"))
    assert n2 == f.shape[0]
    s.close()
    t.close()
