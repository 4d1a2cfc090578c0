"""Seeded input sets for the eight interaction kernels.

Shared by the golden-vector generator (tests/golden/make_golden.py), the oracle tests and the GPU
parity tests, so every check sees the same inputs.  Sinks are a leaf-sized cluster; sources are
either nearby particles (PP), or multipoles built by an FP64 P2M of seeded particle clusters and
rounded to float (PC/CC), mirroring what the tree walk hands the kernels.
"""
import zlib

import numpy as np

MULTI_DTYPE = np.dtype([("m", np.float32, (17,)), ("x", np.int32, (3,))])

# the two kernel-parameter sets (options::ewald, soft_len = 0.02 N^-1/3, m = 1/N)
PARAMS = {
    "periodic": {"ewald": 1, "soft_len": 0.02 * 32768 ** (-1.0 / 3.0), "m": 1.0 / 32768},
    "open": {"ewald": 0, "soft_len": 0.02 * 2097152 ** (-1.0 / 3.0), "m": 1.0 / 2097152},
}


def double_to_pos(d):
    """tigergrav/position.hpp:44-51: lround((d - 0.5) 2^32), wrapped into int32."""
    z = np.rint((np.asarray(d, np.float64) - 0.5) * 4294967296.0).astype(np.int64)
    z = np.where(z >= 2 ** 31, z - 2 ** 32, z)
    return z.astype(np.int32)


def pos_to_double(p):
    return np.asarray(p, np.float64) * (1.0 / 4294967296.0) + 0.5


def _sym_index_lists():
    m2 = [(a, b) for a in range(3) for b in range(a, 3)]
    m3 = [(a, b, c) for a in range(3) for b in range(a, 3) for c in range(b, 3)]
    return m2, m3


def p2m(xd, m, com_pos):
    """FP64 multipole about the (quantised) COM: [M, M(jk) x6, M(jkl) x10], as float32."""
    com = pos_to_double(com_pos)
    dx = xd - com[None, :]
    m2, m3 = _sym_index_lists()
    out = np.zeros(17, np.float64)
    out[0] = m * len(xd)
    for i, (a, b) in enumerate(m2):
        out[1 + i] = m * np.sum(dx[:, a] * dx[:, b])
    for i, (a, b, c) in enumerate(m3):
        out[7 + i] = m * np.sum(dx[:, a] * dx[:, b] * dx[:, c])
    return out.astype(np.float32)


def make_multipoles(rng, n, center, spread, box_r, m, npart=(4, 64)):
    """n multipoles of seeded clusters whose centres lie in a shell of radius `spread` around
    `center`; each cluster has radius box_r and npart particles of mass m."""
    out = np.zeros(n, MULTI_DTYPE)
    for i in range(n):
        v = rng.normal(size=3)
        c = np.asarray(center) + spread * (0.6 + 0.4 * rng.random()) * v / np.linalg.norm(v)
        k = int(rng.integers(npart[0], npart[1] + 1))
        xd = (c[None, :] + box_r * (rng.random((k, 3)) - 0.5)) % 1.0
        com = double_to_pos(xd.mean(axis=0))
        out[i]["m"] = p2m(xd, m, com)
        out[i]["x"] = com
    return out


def cluster(rng, n, center, r):
    return double_to_pos((np.asarray(center)[None, :] + r * (rng.random((n, 3)) - 0.5)) % 1.0)


def kernel_cases(kind, params, seed=1234):
    """List of cases {name, x, y, f0 | L0, do_phi} for one kernel kind."""
    rng = np.random.default_rng(seed + zlib.crc32(kind.encode()) % 1000)
    h, m = params["soft_len"], params["m"]
    c0 = np.array([0.3, 0.6, 0.45])
    cases = []
    sink_part = kind.startswith("PP") or kind.startswith("PC")
    ewald = kind.endswith("ewald")
    if sink_part:
        sizes_s = [1, 17, 64]
        sizes_j = [1, 15, 16, 17, 200] if ewald else [1, 15, 16, 17, 1000]
    else:
        sizes_s = [1]
        sizes_j = [1, 15, 16, 17, 100] if ewald else [1, 15, 16, 17, 300]
    for ns in sizes_s:
        for nj in sizes_j:
            if ewald and sink_part and ns * nj > 64 * 64:
                continue
            x = cluster(rng, ns, c0, 0.03)
            if not sink_part:
                x = x[0]
            if kind.startswith("PP") or kind.startswith("CP"):
                # near sources: neighbouring leaves, plus the sink cluster itself for PP
                y = cluster(rng, nj, c0 + (0.05 if kind.startswith("CP") else 0.0), 0.12)
                if kind.startswith("PP") and sink_part and nj >= ns:
                    y[:ns] = x.reshape(-1, 3)[:ns]  # coincident pairs, as a leaf's own range
            else:
                dist = 0.35 if ewald else 0.2
                y = make_multipoles(rng, nj, c0, dist, 0.04, m)
            case = {"name": f"{kind}_s{ns}_j{nj}", "x": x, "y": y, "do_phi": 1}
            if sink_part:
                case["f0"] = rng.normal(size=(ns, 4))
            else:
                case["L0"] = rng.normal(size=35)
            cases.append(case)
    # special geometries
    if kind == "PP_direct":
        X = double_to_pos(np.array([[0.5, 0.5, 0.5]]))
        hp = h * 4294967296.0
        offs = np.array([0, 1, 0.25 * hp, 0.5 * hp - 1, 0.5 * hp, 0.5 * hp + 1, 0.75 * hp, hp - 1, hp, hp + 1, 1.5 * hp, 3 * hp])
        Y = np.repeat(X, len(offs), axis=0).astype(np.int64)
        Y[:, 0] += np.rint(offs).astype(np.int64)
        Y[3:6, 1] += 7  # off-axis variants near the h/2 switch
        cases.append({"name": "PP_direct_spline_edges", "x": X, "y": Y.astype(np.int32), "do_phi": 1,
                      "f0": np.zeros((1, 4))})
        cases.append({"name": "PP_direct_nophi", "x": cluster(rng, 33, c0, 0.02), "y": cluster(rng, 70, c0, 0.05),
                      "do_phi": 0, "f0": rng.normal(size=(33, 4))})
        cases.append({"name": "PP_direct_empty_source", "x": cluster(rng, 5, c0, 0.02),
                      "y": np.zeros((0, 3), np.int32), "do_phi": 1, "f0": np.zeros((5, 4))})
    if kind == "PP_ewald":
        X = double_to_pos(np.array([[0.25, 0.5, 0.5]]))
        seps = np.array([[0.5, 0, 0], [0.5, 0.5, 0], [0.5, 0.5, 0.5], [-0.5, 0.5, 0], [0, 0, 0], [2 ** -32, 0, 0],
                         [1e-6, 0, 0], [2e-6, 0, 0], [1e-3, 2e-3, -1e-3], [0.1, 0.2, 0.3], [0.49, -0.3, 0.1]])
        Y = double_to_pos((pos_to_double(X)[0][None, :] - seps) % 1.0)
        cases.append({"name": "PP_ewald_symmetry_points", "x": X, "y": Y, "do_phi": 1, "f0": np.zeros((1, 4))})
    if kind == "CP_ewald":
        X = double_to_pos(np.array([0.7, 0.2, 0.9]))
        Y = np.stack([X, X + np.array([3, 0, 0], np.int32), X + np.array([400000, -200000, 7], np.int32)])
        cases.append({"name": "CP_ewald_coincident", "x": X, "y": Y, "do_phi": 1, "L0": np.zeros(35)})
    return cases
