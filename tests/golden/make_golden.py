"""Generate the kernel golden vectors from the compiled reference.

Runs only where the reference source tree exists (it builds oracle/_ref first).  For each of the
two kernel-parameter sets (periodic: ewald=1, N=32768; open: ewald=0, N=2^21) a fresh process sets
the reference's options once -- the reference freezes them at first use, src/gravity.cpp:176-187
-- and runs all eight reference kernels on the seeded cases of tests/kernel_cases.py.  Inputs,
initial accumulators and outputs (plus the reference's flop counts) are written to
tests/golden/kernels_<set>.npz.

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, TESTS)

import oracle_bind as ob  # noqa: E402
from kernel_cases import PARAMS, kernel_cases  # noqa: E402


def generate(cfg):
    P = PARAMS[cfg]
    R = ob.ref_lib()
    R.tgref_set_options(P["ewald"], P["soft_len"], 1.0, int(round(1.0 / P["m"])), 0.7)
    out = {"params": np.array([P["ewald"], P["soft_len"], P["m"]], np.float64)}
    names = []
    for kind in ob.KINDS:
        for c in kernel_cases(kind, P):
            key = c["name"]
            names.append(key)
            out[key + "/x"] = np.asarray(c["x"])
            out[key + "/y"] = np.asarray(c["y"])
            out[key + "/do_phi"] = np.array(c["do_phi"])
            if kind[:2] in ("PP", "PC"):
                res, flop = ob.call_ref(R, kind, c["x"], c["y"], f=c["f0"], do_phi=c["do_phi"])
                out[key + "/f0"] = np.asarray(c["f0"], np.float64)
            else:
                res, flop = ob.call_ref(R, kind, c["x"], c["y"], Lx=c["L0"])
                out[key + "/L0"] = np.asarray(c["L0"], np.float64)
            out[key + "/out"] = res
            out[key + "/flop"] = np.array(flop, np.uint64)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, f"kernels_{cfg}.npz"), **out)
    print(f"{cfg}: {len(names)} cases")


if __name__ == "__main__":
    if len(sys.argv) > 1:
        generate(sys.argv[1])
    else:
        subprocess.run(["make", "-s", "-C", ob.ORACLE_DIR, "ref"], check=True)
        for cfg in PARAMS:  # one process per parameter set: the reference freezes its options
            subprocess.run([sys.executable, os.path.abspath(__file__), cfg], check=True)
