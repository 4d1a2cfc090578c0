"""Autotuner of the batched solve: picks the leaf size, the CC/CP chunk per partial slot, the
PP-direct kernel form (one work group per item, or persistent queue-driven) and the form of the
local-expansion pass (per-leaf path sums, or level by level) on the device time of a few solves of
the caller's own particle set."""
import itertools

from .solver import Solver


def autotune(x, ewald, soft_len, m, theta, leaves=(8, 16, 32, 64, 128), chunks=(32, 64), persistent=(0, 1), sweep=(0, 1),
             solves=3, device=0):
    """Returns (best, table): best = {"leaf", "node_chunk", "pp_persistent", "local_sweep", "ms"};
    table lists every configuration tried with its mean device ms per solve (sum over the solve's
    kernels)."""
    table = []
    for leaf, chunk in itertools.product(leaves, chunks):
        s = Solver(ewald=ewald, soft_len=soft_len, m=m, theta=theta, parts_per_node=leaf, device=device)
        s.tune("node_chunk", chunk)
        s.set_particles(x)
        for p, w in itertools.product(persistent, sweep):
            s.tune("pp_persistent", p)
            s.tune("local_sweep", w)
            s.solve()  # warm
            s.synchronize()
            s.collect_ms()
            for _ in range(solves):
                s.solve()
            s.synchronize()
            ms, n = s.collect_ms()
            table.append({"leaf": leaf, "node_chunk": chunk, "pp_persistent": p, "local_sweep": w, "ms": sum(ms.values())})
        s.close()
    best = min(table, key=lambda r: r["ms"])
    return dict(best), table


def apply(solver, best):
    """Set a tuned configuration's knobs on a solver built with best["leaf"] (before set_particles)."""
    solver.tune("node_chunk", best["node_chunk"])
    solver.tune("pp_persistent", best["pp_persistent"])
    solver.tune("local_sweep", best.get("local_sweep", 0))
