"""The oracle against the compiled reference's golden vectors, bit for bit.

The fixtures (tests/golden/kernels_*.npz) hold the output of the reference's own
src/gravity.cpp, compiled with contraction off (oracle/Makefile), on the seeded cases of
tests/kernel_cases.py, for both kernel-parameter sets.  The oracle restates the same float steps
and the same 16-lane FP64 summation order, so equality is exact (NaN positions included).
"""
import os

import numpy as np
import pytest

import oracle_bind as ob

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(cfg):
    z = np.load(os.path.join(GOLDEN, f"kernels_{cfg}.npz"), allow_pickle=False)
    ew, h, m = z["params"]
    return z, {"ewald": int(ew), "soft_len": float(h), "m": float(m)}


@pytest.mark.parametrize("cfg", ["periodic", "open"])
@pytest.mark.parametrize("kind", ob.KINDS)
def test_oracle_matches_reference_bitwise(oracle, cfg, kind):
    z, P = load(cfg)
    names = [str(n) for n in z["names"] if str(n).startswith(kind)]
    assert names, "no golden cases"
    for name in names:
        x, y = z[name + "/x"], z[name + "/y"]
        if kind[:2] in ("PP", "PC"):
            if kind[:2] == "PC":
                y = y.view(ob.MULTI_DTYPE) if y.dtype != ob.MULTI_DTYPE else y
            out, flop = ob.call_kernel("tgo_", kind, P, x, y, f=z[name + "/f0"], do_phi=int(z[name + "/do_phi"]))
        else:
            out, flop = ob.call_kernel("tgo_", kind, P, x, y, L=z[name + "/L0"])
        ref = z[name + "/out"]
        assert flop == int(z[name + "/flop"]), name
        same = (out == ref) | (np.isnan(out) & np.isnan(ref))
        assert same.all(), f"{name}: {np.sum(~same)} of {out.size} differ, max |d| {np.nanmax(np.abs(out - ref))}"
