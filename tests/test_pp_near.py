"""The PP-direct work items of the batched solve (host side, no GPU needed).

Each sink leaf's PP source list is reordered so that the sources whose leaf box comes within the
softening length of the sink leaf come first; the item records that prefix's length and the GPU
kernel skips the softened branch on every tile past it.  The reordering must therefore
  * keep the list a permutation of the walk's own PP list for that leaf (src/tree.cpp:584-602), and
  * put every source within h of any of the item's sinks inside the prefix, brute-force checked
    over all pairs (minimum image when periodic), on inputs with many pairs closer than h.
"""
import numpy as np
import pytest

from tigergrav_amd import ics
from tigergrav_amd.solver import Solver

COUNT_MASK, NEAR_SHIFT, NEAR_ALL = 127, 7, (1 << 25) - 1


def _solver(x, ewald, h, ppn):
    n = x.shape[0]
    s = Solver(ewald=ewald, soft_len=h, m=1.0 / n, theta=0.7, parts_per_node=ppn, device=-1)
    s.set_particles(x)
    return s


def _expand(ranges):
    return np.concatenate([np.arange(b, e) for b, e in ranges.reshape(-1, 2)]) if len(ranges) else np.zeros(0, int)


CASES = [
    ("uniform_periodic", lambda: ics.uniform(4096, seed=11), True, 32),
    ("plummer_open", lambda: ics.plummer_spheres(8192, seed=12), False, 64),
    ("clumped_periodic", None, True, 32),
]


def _clumped(seed=13):
    # 64 tight clumps straddling the periodic faces: many pairs closer than h across the boundary
    rng = np.random.default_rng(seed)
    c = rng.random((64, 3))
    c[:8, 0] = 0.9995
    x = (c[:, None, :] + 2e-3 * rng.standard_normal((64, 64, 3))).reshape(-1, 3) % 1.0
    return ics.double_to_pos(x)


@pytest.mark.parametrize("name,make,ewald,ppn", CASES, ids=[c[0] for c in CASES])
def test_near_prefix_holds_every_close_pair(name, make, ewald, ppn):
    x = _clumped() if make is None else make()
    n = x.shape[0]
    h = 2e-3 if make is None else 4.0 * ics.soft_len(n)  # large softening: many close pairs
    s = _solver(x, ewald, h, ppn)
    pos = s.export("pos").astype(np.float64) * 2.0 ** -32
    items = s.export("pp_d_items")
    lst = s.export("pp_d_list")
    ptr = s.export("pp_d_ptr")
    ranges = s.export("pp_d")
    rng_ = s.export("node_range")
    leaves = s.export("leaves")
    leaf_of = {int(rng_[lf, 0]): int(lf) for lf in leaves if rng_[lf, 1] > rng_[lf, 0]}
    close_pairs = 0
    for it in items:
        first, y, lb, le = (int(v) for v in it)
        cnt, near = y & COUNT_MASK, (y & 0xFFFFFFFF) >> NEAR_SHIFT
        assert 1 <= cnt <= 64 and near != NEAR_ALL and 0 <= near <= le - lb
        # the whole leaf's list is a permutation of the walk's PP list
        lf = max(b for b in leaf_of if b <= first)
        lf = leaf_of[lf]
        src = lst[lb:le]
        want = _expand(ranges[ptr[lf]:ptr[lf + 1]])
        assert np.array_equal(np.sort(src), np.sort(want))
        d = pos[first:first + cnt, None, :] - pos[src][None, :, :]
        if ewald:
            d -= np.round(d)
        r = np.sqrt((d * d).sum(-1))
        close = np.nonzero((r <= h * (1 + 1e-6)).any(axis=0))[0]
        close_pairs += len(close)
        assert close.size == 0 or close.max() < near, f"close source at list offset {close.max()} past prefix {near}"
    assert close_pairs > len(items)  # the inputs exercise the prefix beyond each leaf's own sinks
