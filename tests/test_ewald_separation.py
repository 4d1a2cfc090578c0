"""tg_ewald_near_separation / tg_ewald_far_separation (src/gravity.cpp:10-23) through the C ABI, no
GPU needed: the minimum-image distance sqrt(sum_d min(|x_d|, 1 - |x_d|)^2) and max(0.25, it).

Pinned by what the geometry fixes: periodic images (x and x + n give the same distance), the
distance to the nearest lattice point by brute force over the 27 neighbouring images, the values at
the symmetry points of the cell, and the far separation's floor of 0.25 (the radius argument is
unused, as in the reference)."""
import itertools

import numpy as np


def _fns():
    from tigergrav_amd import gravity as tg

    return tg.ewald_near_separation, tg.ewald_far_separation


def test_near_separation_is_the_minimum_image_distance():
    near, _ = _fns()
    rng = np.random.default_rng(21)
    for x in rng.uniform(-1.0, 1.0, (300, 3)):
        brute = min(np.linalg.norm(x - np.array(n)) for n in itertools.product((-2, -1, 0, 1, 2), repeat=3))
        assert abs(near(x) - brute) < 1e-14


def test_near_separation_symmetry_points_and_periodicity():
    near, _ = _fns()
    assert near([0.0, 0.0, 0.0]) == 0.0
    assert near([0.5, 0.0, 0.0]) == 0.5
    assert abs(near([0.5, 0.5, 0.5]) - np.sqrt(0.75)) < 1e-15
    assert abs(near([0.25, -0.75, 0.0]) - np.sqrt(0.125)) < 1e-15
    rng = np.random.default_rng(22)
    for x in rng.uniform(-0.5, 0.5, (50, 3)):
        for n in ((1, 0, 0), (0, -1, 1)):
            # x + n rounds, so the distances agree to rounding only
            assert abs(near(x) - near(x + np.array(n))) < 1e-15


def test_far_separation_floor():
    _, far = _fns()
    assert far([0.1, 0.0, 0.0], 0.3) == 0.25
    assert far([0.0, 0.0, 0.0]) == 0.25
    assert far([0.4, 0.0, 0.0], 99.0) == 0.4
    assert abs(far([0.5, 0.5, 0.0]) - np.sqrt(0.5)) < 1e-15
