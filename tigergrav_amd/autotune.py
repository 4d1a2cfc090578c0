"""Autotuner of the batched solve: picks the leaf size, the CC/CP chunk per partial slot (chunks
longer than 64 entries are one wave each, its lanes adding several entries), the PP-direct kernel
form (one work group per item, or persistent queue-driven), the form of the local-expansion pass
(per-leaf path sums, or level by level) and whether graph replays run the node kinds on a second
stream.  The criterion is the time per solve of graph replays of the caller's own particle set
(what the benchmark measures: with the node kinds beside the leaf kinds, so the sum of per-kind
kernel times would misjudge it), among the leaf sizes an optional acceptance test passes (the
leaf size and the opening angle set the force error; the other knobs do not)."""
import itertools
import time

from .solver import Solver


def autotune(x, ewald, soft_len, m, theta, leaves=(8, 16, 32, 64, 128), chunks=(256, 4096), persistent=(0, 1), sweep=(0, 1),
             overlap=(1, 0, 2), solves=10, device=0, accept=None):
    """Returns (best, table): best = {"leaf", "node_chunk", "pp_persistent", "local_sweep", "overlap", "ms"};
    table lists every configuration tried with its ms per solve over `solves` graph replays
    (host clock, the device synchronised before and after).

    accept, if given, is called once per leaf size with a solver of that leaf size after one solve
    and returns (ok, info): a leaf size it rejects (an accuracy bound the forces miss, say; the
    other knobs change the forces only by rounding) is recorded with its info and not timed.  The
    best configuration is the fastest accepted one; with none accepted, the fastest of all, marked
    "accepted": False."""
    table = []
    verdict = {}
    for leaf, chunk in itertools.product(leaves, chunks):
        if leaf in verdict and not verdict[leaf][0]:
            continue
        s = Solver(ewald=ewald, soft_len=soft_len, m=m, theta=theta, parts_per_node=leaf, device=device)
        s.tune("node_chunk", chunk)
        s.set_particles(x)
        if accept is not None and leaf not in verdict:
            s.solve()
            s.synchronize()
            verdict[leaf] = accept(s)
            if not verdict[leaf][0]:
                table.append({"leaf": leaf, "rejected": verdict[leaf][1]})
                s.close()
                continue
        for p, w, o in itertools.product(persistent, sweep, overlap):
            s.tune("pp_persistent", p)
            s.tune("local_sweep", w)
            s.tune("overlap", o)
            for _ in range(3):  # capture and warm
                s.solve_graph()
            s.synchronize()
            t0 = time.perf_counter()
            for _ in range(solves):
                s.solve_graph()
            s.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / solves
            row = {"leaf": leaf, "node_chunk": chunk, "pp_persistent": p, "local_sweep": w, "overlap": o, "ms": ms}
            if leaf in verdict:
                row["accepted"] = verdict[leaf][1]
            table.append(row)
        s.close()
    timed = [r for r in table if "ms" in r]
    if not timed:  # every leaf size rejected: time them all and mark the choice
        best, table2 = autotune(x, ewald, soft_len, m, theta, leaves, chunks, persistent, sweep, overlap, solves, device)
        best["accepted"] = False
        return best, table + table2
    best = dict(min(timed, key=lambda r: r["ms"]))
    return best, table


def apply(solver, best):
    """Set a tuned configuration's knobs on a solver built with best["leaf"] (before set_particles);
    knobs missing from `best` keep their defaults."""
    for knob in ("node_chunk", "pp_persistent", "local_sweep", "overlap"):
        if best.get(knob) is not None:
            solver.tune(knob, best[knob])
