"""Benchmark: periodic force solve of 2,097,152 particles on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload zeldovich128] [--theta 0.3] [--leaf 0] [--no-graph]

Workload (default): 128^3 particles displaced by a Zel'dovich field of a seeded Gaussian random
field with P(k) ~ k^-2 (cut at the grid Nyquist), rms displacement 0.3 cell, periodic (Ewald),
softening 0.02 N^(-1/3), equal masses 1/N.  The timed solve runs at theta = 0.3: the largest
opening angle whose force error meets RMS |da|/|a| <= 1e-3 against the FP64 direct sum on this
workload (the `accuracy` field shows it).  The leaf size (8/16/32/64/128), the CC/CP chunk, the PP
kernel form, the form of the local pass and the stream overlap are autotuned on this workload's
own graph-replay time, among the leaf sizes whose error on the same FP64 sample meets the bound
(the leaf size moves the error too: smaller leaves turn near pairs into multipole interactions);
--leaf N with --node-chunk/--pp-persistent/--local-sweep/--overlap fixes them instead.  The timed
steps replay the solve from a captured HIP graph.  A second leg, `theta07`, times the same particles at
theta = 0.7 (the workload's specified angle; about 10x faster, with RMS error ~0.15, which does not
meet 1e-3).

One step = one force solve on the device over the resident tree and interaction lists: the CC/CP
node kernels, the local-expansion pass (reduce + L2L + L2P) and the PC/PP leaf kernels.  The tree,
multipoles, lists and the kernels' work arrays are built on the device before the timed region
(`build_s`: the first build, and the median of three rebuilds on the same solver, the steady
state of a time step; `s_per_force_solve_end_to_end` is that rebuild plus one solve, host
particles to forces).  Per-kind kernel times (and the
roofline) come from HIP events on the solver's stream around each kind's kernels, over K eager
solves run right after the graph-replayed ones (a captured graph carries no timers).
`profile_cmd` is the command that reruns the timed configuration with its knobs fixed, under
which the committed profiler summaries of the dominant kernel were taken.

Prints one JSON line.  `value` is interactions per second (the reference's interaction census,
src/tree.cpp:483-494,618-628, divided by the device time of one solve).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "interactions/s + s per force solve, 2M periodic, RMS |Δa|/|a| ≤1e-3"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# algorithmic flops per interaction (add/sub/mul/div/sqrt/min/max = 1, FMA = 2), counted from the
# reference formulas.  PP_direct: the far-field branch (r >= h) of src/gravity.cpp:231-287 with the
# potential -- dX 3, r2 5, sqrt 1, 1/max(r, h/2) 2, rinv^3 2, g -= dX*M*rinv^3 9, phi -= M*rinv 2 --
# which every pair of the 2M workloads but a few in 1e6 takes; the spline branch for r < h adds 52
# and the reference's SIMD code evaluates all three branches for every pair (76 in total).
ALG_FLOPS = {"PP_direct": 24, "PC_direct": 478, "CC_direct": 621, "CP_direct": 321, "PP_ewald": 30057,
             "PC_ewald": 127958, "CC_ewald": 128101, "CP_ewald": 127801}
KINDS = ("PP_direct", "PC_direct", "CC_direct", "CP_direct", "PP_ewald", "PC_ewald", "CC_ewald", "CP_ewald")


def make_workload(name):
    from tigergrav_amd import ics

    if name == "zeldovich128":
        return ics.zeldovich(128, seed=3, n_s=-2.0, rms_cells=0.3), True, (
            "zeldovich 128^3 = 2,097,152 particles, P(k)~k^-2 cut at grid Nyquist, rms displacement 0.3 cell, "
            "seed 3, periodic Ewald")
    if name == "zeldovich64":
        return ics.zeldovich(64, seed=4, n_s=-2.0, rms_cells=0.3), True, (
            "zeldovich 64^3 = 262,144 particles, P(k)~k^-2 cut at grid Nyquist, rms displacement 0.3 cell, "
            "seed 4, periodic Ewald")
    if name == "zeldovich256":
        return ics.zeldovich(256, seed=5, n_s=-1.5, rms_cells=0.5), True, (
            "zeldovich 256^3 = 16,777,216 particles, P(k)~k^-1.5 cut at grid Nyquist, rms displacement 0.5 cell, "
            "seed 5, periodic Ewald")
    if name == "plummer2M":
        return ics.plummer_spheres(2 ** 21, seed=2), False, (
            "8 Plummer spheres (a = 0.02) + 20% uniform background, 2,097,152 particles, seed 2, open boundaries")
    if name == "uniform32k":
        return ics.uniform(32768, seed=1), True, "uniform 32,768 particles, seed 1, periodic Ewald"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(s, ewald, h, m, theta, leaf, target_s=15.0):
    """Time the reference's own compiled kernels (oracle/_ref, if built) -- else the oracle's C
    restatement -- over a strided sample of the same interaction lists, on all host threads.
    Returns the baseline record and, when the sample is the whole solve, its forces (input order)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    T = s.tree_arrays()
    N = T["pos"].shape[0]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    kind = "reference" if ob.ref_available() else "port"
    if kind == "reference":
        R = ob.ref_lib()
        R.tgref_set_options(int(ewald), h, 1.0, N, theta)
        run = lambda stride: ob.ref_solve(R, T, stride)
    else:
        run = lambda stride: ob.oracle_solve({"ewald": int(ewald), "soft_len": h, "m": m}, T, stride)
    # interactions in the sample (census restricted to sampled nodes / leaves)
    nr = T["node_range"]
    ns = (nr[:, 1] - nr[:, 0]).astype(np.int64)

    def sample_count(stride):
        tot = 0
        nodes = np.arange(len(nr))
        sel = nodes[nodes % stride == 0]
        for k in ("cc_d", "cc_e"):
            p = T[k + "_ptr"]
            tot += int(np.sum(p[sel + 1] - p[sel]))
        for k in ("cp_d", "cp_e"):
            p, v = T[k + "_ptr"], T[k].reshape(-1, 2)
            lens = v[:, 1] - v[:, 0]
            cs = np.concatenate([[0], np.cumsum(lens)])
            tot += int(np.sum(cs[p[sel + 1] // 2] - cs[p[sel] // 2]))
        leaves = sel[T["node_child"][sel, 0] < 0]
        for k in ("pc_d", "pc_e"):
            p = T[k + "_ptr"]
            tot += int(np.sum((p[leaves + 1] - p[leaves]) * ns[leaves]))
        for k in ("pp_d", "pp_e"):
            p, v = T[k + "_ptr"], T[k].reshape(-1, 2)
            lens = v[:, 1] - v[:, 0]
            cs = np.concatenate([[0], np.cumsum(lens)])
            tot += int(np.sum((cs[p[leaves + 1] // 2] - cs[p[leaves] // 2]) * ns[leaves]))
        return tot

    # a probe on every 64th node's lists, then the largest sample (the whole solve if it fits) that
    # takes about target_s seconds of wall time on all threads
    stride = 64
    t0 = time.time()
    run(stride)
    dt = time.time() - t0
    stride = max(1, min(64, math.ceil(stride * dt / target_s)))
    t0 = time.time()
    forces_tree, _ = run(stride)
    dt = time.time() - t0
    cnt = sample_count(stride)
    return {"value": cnt / dt, "unit": "interactions/s", "cores": threads, "kind": kind,
            "sample": ("the whole solve" if stride == 1 else f"every {stride}th node's and leaf's interaction lists of "
                       "the same solve") + f" ({cnt} interactions, {dt:.2f} s wall on {threads} threads)",
            "s_per_force_solve_extrapolated": dt * sum(s.info()["count"].values()) / max(cnt, 1)}, (
        _input_order(forces_tree, T["order"]) if stride == 1 else None)


def _input_order(f_tree, order):
    f = np.empty_like(f_tree)
    f[order] = f_tree
    return f


def device_free_bytes():
    """Free device memory on device 0 (hipMemGetInfo), through the HIP runtime the library uses."""
    import ctypes as C

    from tigergrav_amd._lib import lib

    lib()  # loads the process's one HIP runtime first
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        return 0
    return int(free.value)


def adjoint_leg(s, x, ewald, h, m, forward_ms, steps):
    """The adjoint of the solve (tg_solver_adjoint) on the same particle set: warm device time per
    call against the forward's, a directional central difference of the frozen-tree solve, and the
    difference from the FP64 direct-sum gradient on a 4096-particle subset."""
    from tigergrav_amd.accuracy import direct_gradient

    N = x.shape[0]
    rng = np.random.default_rng(5)
    w = rng.standard_normal((N, 3)).astype(np.float32)
    s.adjoint(w)  # warm
    t0 = time.perf_counter()
    times = []
    for _ in range(max(1, min(steps, 10))):
        g, ms = s.adjoint(w, with_time=True)
        times.append(ms)
    wall = (time.perf_counter() - t0) * 1e3 / len(times)
    adj_ms = float(np.mean(times))
    # directional check on the frozen tree (every particle moved by up to 2^-18 of the box)
    step = rng.integers(-2 ** 14, 2 ** 14, size=x.shape).astype(np.int64)

    def loss(xx):
        s.update_particles(xx)
        s.solve()
        s.synchronize()
        return float(np.sum(w * s.forces()[:, 1:]))

    fd = 0.5 * (loss((x.astype(np.int64) + step).astype(np.int32)) - loss((x.astype(np.int64) - step).astype(np.int32)))
    pred = float(np.sum(g[:, :3] * step * 2.0 ** -32))
    s.update_particles(x)
    sub = np.sort(rng.choice(N, min(N, 4096), replace=False))
    gd = direct_gradient(x, w, sub, ewald, h, m)
    ex = np.linalg.norm(g[sub, :3] - gd[:, :3], axis=1) / np.linalg.norm(gd[:, :3], axis=1)
    return {"ms": adj_ms, "wall_ms_per_call": wall, "forward_ms": forward_ms, "ratio_to_forward": adj_ms / forward_ms,
            "directional_check": {"fd": fd, "adjoint": pred, "rel": abs(fd - pred) / abs(pred)},
            "vs_direct_gradient_rms_rel": float(np.sqrt(np.mean(ex ** 2))), "subset": int(sub.size),
            "loss": "sum_i w_i . g_i, w ~ N(0, I), seed 5"}


def build_solver(x, ewald, h, m, theta, leaf, leaves, knobs, host_walk=False, accept=None):
    """A solver for the particle set at this opening angle, with the leaf size and knobs fixed
    (leaf > 0) or autotuned over `leaves` on graph-replay time (leaf == 0; with accept, among the
    leaf sizes it accepts).  Returns the solver (particles set), the tuned configuration, the
    tuning table and the host build time (s)."""
    from tigergrav_amd.autotune import apply, autotune
    from tigergrav_amd.solver import Solver

    best, table = None, None
    if leaf == 0:
        best, table = autotune(x, ewald, h, m, theta, leaves=leaves, accept=accept)
        leaf = best["leaf"]
    else:
        best = {"leaf": leaf}
        best.update({k: v for k, v in knobs.items() if v is not None})
    s = Solver(ewald=ewald, soft_len=h, m=m, theta=theta, parts_per_node=leaf, device=0)
    if host_walk:  # the host build: tree, multipoles and lists on the host (the same bits)
        s.tune("gpu_walk", 0)
        s.tune("gpu_tree", 0)
    apply(s, best)
    t0 = time.time()
    s.set_particles(x)
    return s, best, table, time.time() - t0


def build_times(info, total):
    """A build's wall time (s) with its parts: tree, multipoles, interaction lists, upload and the
    finer split of the last two (walk set-up, steps, assembly, census; work preparation, transfers)."""
    return {"total": total, "tree": info["t_tree"], "multipoles": info["t_multipoles"], "lists": info["t_walk"],
            "upload": info["t_upload"], "detail": info["t_detail"]}


def time_solves(s, steps, warmup, graph):
    """ms per solve over `steps` back-to-back solves (graph replays, or eager), after `warmup`
    untimed ones; then the per-kind device means over as many eager solves with event timers."""
    step = s.solve_graph if graph else s.solve
    for _ in range(warmup):
        step()
    s.synchronize()
    s.collect_ms()  # drop the warmup solves' event times
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    s.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ms_eager = None
    if graph:
        t0 = time.perf_counter()
        for _ in range(steps):
            s.solve()
        s.synchronize()
        ms_eager = (time.perf_counter() - t0) * 1e3 / steps
    kms, nsolves = s.collect_ms()
    assert nsolves == steps
    return ms, ms_eager, kms


def profile_command(workload, theta, best, steps=3):
    """The bench command that reruns exactly this configuration with every knob fixed (what the
    committed rocprofv3 summaries of the dominant kernel were taken from)."""
    cmd = f"python3 bench.py --workload {workload} --theta {theta} --leaf {best['leaf']}"
    for k, flag in (("node_chunk", "--node-chunk"), ("pp_persistent", "--pp-persistent"), ("local_sweep", "--local-sweep"),
                    ("overlap", "--overlap")):
        if best.get(k) is not None:
            cmd += f" {flag} {best[k]}"
    return cmd + f" --steps {steps} --warmup 1 --no-cpu --no-theta07 --no-step --accuracy-sample 1000"


def roofline_of(workload, theta, best, count, kms):
    """Roofline of the dominant kernel: algorithmic flops per interaction x the interactions of one
    launch / its mean launch time (HIP events on the solver's stream), against the FP32 vector peak;
    HBM bytes per launch from the committed profiler summary of the same configuration."""
    dom = max(KINDS, key=lambda k: kms[k])
    achieved = count[dom] * ALG_FLOPS[dom] / (kms[dom] / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                "flops_per_interaction": ALG_FLOPS[dom], "interactions": count[dom], "kernel_ms": kms[dom]}
    # the ceilings this part sustains for the same instruction mix, measured now on this device
    import ctypes as C
    from tigergrav_amd._lib import lib

    L = lib()
    L.tg_calibrate.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    fma_tf, pair_rate = C.c_double(), C.c_double()
    if L.tg_calibrate(0, 3, C.byref(fma_tf), C.byref(pair_rate)) == 0:
        roofline["measured_fp32_fma_tflops"] = fma_tf.value
        roofline["frac_of_measured_fma"] = achieved / fma_tf.value
        if dom == "PP_direct":
            rate = count[dom] / (kms[dom] / 1e3)
            roofline["pp_pairs_per_s"] = rate
            roofline["pp_pair_loop_ceiling_per_s"] = pair_rate.value
            roofline["frac_of_pair_loop_ceiling"] = rate / pair_rate.value
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        try:
            tr = json.load(open(traffic_file))
            key = f"{workload}:theta{theta:g}:leaf{best['leaf']}:{dom}"
            if key in tr:
                roofline["traffic"] = tr[key]["bytes"]
                roofline["traffic_source"] = tr[key]["source"]
        except Exception:
            pass
    return roofline


def step_leg(s, x, h, m, ewald, reps=10):
    """Kick and drift (part_vect_kick / part_vect_drift, src/part_vect.cpp:341-436,493-540) of the
    workload's particles, device-resident, with the solve's forces: ms per call (median of `reps`
    after a warm call; each call returns its statistics to the host) and the rate at the
    algorithmic bytes per particle -- kick: forces 32 + velocity 12 + rung 1 read, velocity 12 +
    rung 1 written; drift: position 12 + velocity 12 read, position 12 written."""
    import torch

    from tigergrav_amd import step

    n = x.shape[0]
    dev = torch.device("cuda:0")
    f = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    L = step._bind()
    step._check(L.tg_solver_get_forces_device(s._h, f.data_ptr()), "tg_solver_get_forces_device")
    v = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    rung = torch.zeros(n, dtype=torch.int8, device=dev)
    xd = torch.as_tensor(np.ascontiguousarray(x, np.int32)).to(dev)
    P = step.params(n, soft_len=h, ewald=ewald, m=m)
    out = {}
    for name, call, nbytes in (("kick", lambda: step.kick(P, f, v, rung, 0, True), 58),
                               ("drift", lambda: step.drift(P, xd, v, 6, drift_dt=step.rung_to_dt(1.0, 6)), 36)):
        call()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            call()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        gbs = n * nbytes / (ms / 1e3) / 1e9
        out[name] = {"ms": ms, "bytes_per_particle": nbytes, "GB_per_s": gbs, "frac_of_hbm": gbs / PEAK_HBM_GBS}
    out["note"] = "wall time per call including the statistics' return to the host"
    return out


def fof_leg(b=0.2):
    """Friends-of-friends groups (tree::find_groups) at linking length b on a clustered set of the
    same size (the near-lattice workload has almost no pairs within 0.2 of the mean spacing): 2M
    particles in 8 Plummer spheres and a uniform background, periodic, leaf 32.  The first call
    (with the host copy of the device tree it needs) and a repeat."""
    from tigergrav_amd import ics
    from tigergrav_amd.solver import Solver

    x = ics.plummer_spheres(2 ** 21, seed=2)
    n = x.shape[0]
    s = Solver(ewald=True, soft_len=ics.soft_len(n), m=1.0 / n, theta=0.7, parts_per_node=32, device=0)
    s.set_particles(x)
    t0 = time.perf_counter()
    g, ng = s.find_groups(b)
    t1 = time.perf_counter()
    g2, ng2 = s.find_groups(b)
    t2 = time.perf_counter()
    assert ng == ng2 and np.array_equal(g, g2)
    s.close()
    return {"workload": "8 Plummer spheres + background, 2,097,152 particles", "b": b, "groups": int(ng),
            "linked_particles": int(np.sum(g != 0xFFFFFFFFFFFFFF)), "s_first": t1 - t0, "s_repeat": t2 - t1}


def accuracy_of(f, sinks, fref, ns):
    from tigergrav_amd.accuracy import error_stats

    acc = error_stats(f[sinks], fref)
    acc.update({"meets_1e-3": bool(acc["rms"] <= 1e-3), "sample": f"{ns} sinks (seed 7)"})
    return acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="zeldovich128")
    ap.add_argument("--theta", type=float, default=0.3,
                    help="opening angle of the timed solve (0.3: the largest that meets RMS 1e-3 on the 2M workload)")
    ap.add_argument("--leaf", type=int, default=0,
                    help="particles per leaf; 0 autotunes leaf (8, 16, 32, 64, 128), CC/CP chunk, PP kernel form and "
                         "local pass")
    ap.add_argument("--wide-autotune", action="store_true", help="autotune over leaves of 4, 12, 24 and 48 as well")
    ap.add_argument("--max-rms", type=float, default=1e-3,
                    help="the autotuner takes only leaf sizes whose RMS |da|/|a| on the accuracy sample is at most this "
                         "(the metric's bound; 0 takes every leaf size)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-step", action="store_true", help="skip the kick/drift and friends-of-friends legs")
    ap.add_argument("--pp-persistent", type=int, default=None, choices=(0, 1),
                    help="with a fixed --leaf: the persistent (1) or one-group-per-item (0) PP kernel")
    ap.add_argument("--node-chunk", type=int, default=None, help="with a fixed --leaf: CC/CP list entries per partial slot")
    ap.add_argument("--local-sweep", type=int, default=None, choices=(0, 1),
                    help="with a fixed --leaf: the local-expansion pass level by level (1) or by per-leaf path sums (0)")
    ap.add_argument("--overlap", type=int, default=None, choices=(0, 1, 2),
                    help="with a fixed --leaf: graph replays run on one stream (0), the node kinds on a second stream (1), "
                         "and also the Ewald node kinds on a third (2)")
    ap.add_argument("--host-walk", action="store_true",
                    help="build the tree, multipoles and interaction lists on the host instead of the device")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager solves (with per-kind event timers) instead of HIP-graph replays")
    ap.add_argument("--no-theta07", action="store_true",
                    help="skip the theta 0.7 leg (the 2M workload's specified opening angle, which does not meet 1e-3)")
    ap.add_argument("--accuracy-sample", type=int, default=10000,
                    help="sinks checked against the FP64 direct sum (all of them if >= N)")
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the adjoint (gradients of sum_i w_i . g_i, w ~ N(0, I), seed 5) against the forward")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run only a plain device-to-device copy of this many bytes (5 times): the reference "
                         "point that scales the HBM byte counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              "decomposition is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.copy_bytes > 0:
        import torch

        n = args.copy_bytes // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 5}), flush=True)
        return

    from tigergrav_amd import ics
    from tigergrav_amd.accuracy import direct_forces

    x, ewald, desc = make_workload(args.workload)
    N = x.shape[0]
    h, m = ics.soft_len(N), 1.0 / N
    # the adjoint handles leaves of at most 64 particles
    wide = (4, 12, 24, 48) if args.wide_autotune else ()
    leaves = tuple(sorted(l for l in (8, 16, 32, 64, 128) + wide if l <= 64 or not args.adjoint))
    knobs = {"node_chunk": args.node_chunk, "pp_persistent": args.pp_persistent, "local_sweep": args.local_sweep,
             "overlap": args.overlap}
    # the accuracy reference: FP64 direct sum (converged Ewald when periodic) on the GPU for a seeded
    # sample of sinks; the autotuner keeps the leaf sizes that meet the bound on it
    ns = min(N, args.accuracy_sample)
    sinks = np.sort(np.random.default_rng(7).choice(N, ns, replace=False)) if ns < N else np.arange(N)
    t0 = time.perf_counter()
    fref = direct_forces(x, sinks, ewald, h, m)
    t_ref = time.perf_counter() - t0
    accept = None
    if args.max_rms > 0:
        def accept(solver):
            rms = accuracy_of(solver.forces(), sinks, fref, ns)["rms"]
            return rms <= args.max_rms, {"rms": rms, "bound": args.max_rms}
    free0 = device_free_bytes()
    s, best, tuning, t_host = build_solver(x, ewald, h, m, args.theta, args.leaf, leaves, knobs, args.host_walk, accept)
    info = s.info()
    count = info["count"]
    total = sum(count.values())
    graph = not args.no_graph
    # timed region: K solves back to back on the resident data (graph replays by default)
    ms, ms_eager, kms = time_solves(s, args.steps, args.warmup, graph)
    value = total / (ms / 1e3)

    out = {
        "metric": METRIC, "value": value, "unit": "interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "N": N, "theta": args.theta, "leaf": best["leaf"], "ewald": ewald, "softening": h,
                   "pp_persistent": best.get("pp_persistent"), "node_chunk": best.get("node_chunk"),
                   "local_sweep": best.get("local_sweep"), "overlap": best.get("overlap"), "accumulation": "f64",
                   "parallelism": "single GPU"},
        "s_per_force_solve": ms / 1e3,
        "step": "eager solves" if args.no_graph else "HIP-graph replays",
        "ms_per_step_eager": ms_eager,
        "s_per_force_solve_end_to_end": None,  # filled below from the rebuilds
        "build_s": None,
        "interactions": count, "kernel_ms": kms,
        "tuning": {"best": best, "table": tuning} if tuning else None,
        "roofline": roofline_of(args.workload, args.theta, best, count, kms),
        "profile_cmd": profile_command(args.workload, args.theta, best),
    }
    # force accuracy of the timed solve against the reference above
    f = s.forces()
    # end to end: host particles to forces.  The first build on a fresh solver allocates its device
    # buffers; a time step rebuilds on the same solver, so the steady state is the median of three
    # rebuilds (same particles, same lists) plus one solve
    out["build_s"] = {"first": build_times(info, t_host), "lists_on": "host" if args.host_walk else "device",
                      "tree_on": "host" if args.host_walk else "device"}
    reb = []
    for _ in range(3):
        t0 = time.perf_counter()
        s.set_particles(x)
        s.synchronize()
        reb.append(build_times(s.info(), time.perf_counter() - t0))
    reb.sort(key=lambda r: r["total"])
    out["build_s"]["rebuild"] = reb[1]
    out["s_per_force_solve_end_to_end"] = reb[1]["total"] + ms / 1e3
    out["accuracy"] = accuracy_of(f, sinks, fref, ns)
    out["accuracy"].update({"reference": "FP64 direct sum over all N sources" + (", converged Ewald sum" if ewald else ""),
                            "s": t_ref})
    # device memory held by the solver (all of it stays allocated between solves): resident data
    # after the forward, and after the adjoint's buffers are added
    fp = {"forward_bytes": free0 - device_free_bytes()}
    if args.adjoint:
        out["adjoint"] = adjoint_leg(s, x, ewald, h, m, ms, args.steps) if best["leaf"] <= 64 else {
            "skipped": "the adjoint handles leaves of at most 64 particles"}
        fp["forward_plus_adjoint_bytes"] = free0 - device_free_bytes()
    fp["method"] = "hipMemGetInfo difference from before the solver was created (the accuracy checks free their buffers)"
    out["hbm_footprint"] = fp
    if not args.no_step:
        out["step"] = step_leg(s, x, h, m, ewald)
        out["fof"] = fof_leg()
    if not args.no_cpu:
        out["cpu_baseline"], f_cpu = cpu_baseline(s, ewald, h, m, args.theta, best["leaf"])
        if f_cpu is not None:
            # the same solve done on the host by the CPU baseline: its error on the same sample
            out["cpu_baseline"]["accuracy_rms"] = accuracy_of(f_cpu, sinks, fref, ns)["rms"]
    s.close()
    if not args.no_theta07 and args.workload == "zeldovich128" and abs(args.theta - 0.7) > 1e-9:
        # the workload's specified opening angle, 0.7: faster, but its error is the reference
        # algorithm's own truncation on near-lattice forces, far above 1e-3 (autotuned on the
        # workload's leaf set 32/64/128, timed the same way, same accuracy sample)
        s7, best7, _, t7 = build_solver(x, ewald, h, m, 0.7, 0, (32, 64, 128), knobs)
        ms7, _, kms7 = time_solves(s7, args.steps, args.warmup, graph)
        c7 = s7.info()["count"]
        out["theta07"] = {"theta": 0.7, "best": best7, "ms_per_solve": ms7,
                          "interactions_per_s": sum(c7.values()) / (ms7 / 1e3), "kernel_ms": kms7,
                          "s_per_force_solve_end_to_end": t7 + ms7 / 1e3,
                          "accuracy": accuracy_of(s7.forces(), sinks, fref, ns)}
        s7.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
