"""Autotuner of the engine's launch shapes; writes the tuning table that
liblouvain_amd.so reads at lv_load_edges (tuning.txt beside the library).

The reference fixes its launch shapes (512-thread blocks, 32-lane rows,
LouvainCUDA/OneLevelKernel.cu:520,759).  Here three knobs set how the skewed
rows and coarsening segments are cut (lv_config.hub_q, hub_part_target,
chunk_mul); results do not depend on them, time does.  For each size class the
tuner times every candidate on a sample graph of that class (device-resident
runs, median of --reps after one warmup), checks that every candidate's final
partition equals the default's, and writes one table row per class:

    max_nnz hub_q hub_part_target chunk_mul

Run on the GPU it is meant for (one minute or two):
    python -m louvaingpu_amd.tune --out louvaingpu_amd/tuning.txt
"""
from __future__ import annotations

import argparse
import datetime
import itertools
import sys
import time

import numpy as np

from . import Engine, generators

HUB_Q = (512, 1024, 2048, 4096)
PART_TARGET = (1024, 2048, 4096)
CHUNK_MUL = (8, 16, 32, 64)
DEFAULT = {"hub_q": 4096, "hub_part_target": 2048, "chunk_mul": 32}  # the built-in shapes
MARGIN = 0.02  # a candidate replaces the default only when this much faster

# (size class upper bound in directed entries, sample graph)
CLASSES = (
    (1 << 21, "sbm20000", lambda: generators.sbm(20000, 100, 0.05, 0.001, 5)),
    (1 << 23, "rmat18", lambda: generators.rmat(18, 16, 0.57, 0.19, 0.19, 5, "int")),
    (1 << 26, "rmat20", lambda: generators.rmat(20, 16, 0.57, 0.19, 0.19, 5, "int")),
    (1 << 29, "rmat22", lambda: generators.rmat(22, 16, 0.57, 0.19, 0.19, 5, "int")),
    (1 << 62, "rmat24", lambda: generators.rmat(24, 16, 0.57, 0.19, 0.19, 5, "int")),
)


def time_config(graph, reps, **knobs):
    """Median device-resident run time (ms) and the final partition."""
    with Engine(**knobs) as eng:
        eng.load_edges(*graph)
        eng.run()  # warmup (arena, code objects)
        ts = []
        for _ in range(reps):
            eng.run()
            ts.append(eng.last_run_ms())
        return float(np.median(ts)), eng.final_assignment()


def tune_class(name, graph, reps, quick):
    grid = list(itertools.product(HUB_Q, PART_TARGET, CHUNK_MUL))
    if quick:
        grid = [(q, t, k) for q, t, k in grid if k in (16, 32) and t != 4096]
    base_ms, base_part = time_config(graph, reps, **DEFAULT)
    best = (base_ms, DEFAULT["hub_q"], DEFAULT["hub_part_target"], DEFAULT["chunk_mul"])
    for q, t, k in grid:
        ms, part = time_config(graph, reps, hub_q=q, hub_part_target=t, chunk_mul=k)
        if not np.array_equal(part, base_part):
            raise SystemExit("tune: hub_q=%d part=%d chunk=%d changed the partition of %s"
                             % (q, t, k, name))
        if ms < best[0]:
            best = (ms, q, t, k)
        print("[tune] %s hub_q=%4d part=%4d chunk=%2d: %.3f ms" % (name, q, t, k, ms),
              flush=True)
    if best[0] > base_ms * (1.0 - MARGIN):  # within timing noise of the default: keep it
        best = (base_ms, DEFAULT["hub_q"], DEFAULT["hub_part_target"], DEFAULT["chunk_mul"])
    return base_ms, best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="louvaingpu_amd/tuning.txt")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="a reduced grid")
    args = ap.parse_args(argv)
    lines = ["# launch-shape tuning table of liblouvain_amd.so, written by",
             "#   python -m louvaingpu_amd.tune %s" % " ".join(argv if argv is not None
                                                            else sys.argv[1:]),
             "# on %s; per size class: the sample graph, default and best median run ms" %
             datetime.date.today().isoformat(),
             "# max_nnz hub_q hub_part_target chunk_mul"]
    rows = []
    for max_nnz, name, make in CLASSES:
        t0 = time.perf_counter()
        graph = make()
        print("[tune] %s: n=%d e=%d (%.1f s)" % (name, graph[0], len(graph[1]),
                                                 time.perf_counter() - t0), flush=True)
        base_ms, (ms, q, t, k) = tune_class(name, graph, args.reps, args.quick)
        lines.append("# %s (nnz %d): default %.3f ms, best %.3f ms" % (name, 2 * len(graph[1]),
                                                                     base_ms, ms))
        rows.append("%d %d %d %d" % (max_nnz, q, t, k))
    with open(args.out, "w") as f:
        f.write("\n".join(lines + rows) + "\n")
    print("[tune] wrote %s" % args.out, flush=True)


if __name__ == "__main__":
    main()
