"""Writes rmat26f_oracle.json: the oracle's results on R-MAT scale 26 with
float32 weights U[1,10] (edge factor 16, a/b/c/d = .57/.19/.19/.05, seed 2,
about 1.05 G edges and 2.1 G directed CSR entries).

The oracle needs minutes and >100 GB of host memory at this size, so the
full-size GPU test compares the engine with these digests instead of running
the oracle itself.  Recorded: the generator digest of the edge list, per level
n, k, m (units), passes, the Q of every break test, the returned Q and lower
clamp, SHA-256 of the level's renumbered assignment and of its Partition
arrays, and SHA-256 of the final flattened partition with its community count
and FP64 modularity.

Run from the repository root (a host with ~200 GB of memory):
    ORACLE_PROGRESS=1 python tests/golden/make_rmat26f_fixture.py [out.json]
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from louvaingpu_amd import generators  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rmat26f_oracle.json")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else OUT
    t0 = time.perf_counter()
    n, src, dst, w = generators.rmat26_float()
    gen_s = time.perf_counter() - t0
    print("[fixture] generated n=%d e=%d in %.1f s" % (n, len(src), gen_s), flush=True)
    gdig = sha(np.int64(n), src, dst, w)
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    r = oracle.run(n, src, dst, w, threads=threads)
    run_s = time.perf_counter() - t0
    print("[fixture] oracle %d levels in %.1f s (%d threads)" % (r.num_levels, run_s, threads),
          flush=True)
    levels = []
    for l in range(r.num_levels):
        node, comm = r.partition(l)
        levels.append({
            "n": r.level_n(l), "k": r.level_k(l), "m_units": r.level_m(l),
            "passes": r.level_passes(l),
            "pass_q": [r.pass_q(l, p).hex() for p in range(1, r.level_passes(l) + 1)],
            "q": r.level_q(l).hex(), "lower": r.level_lower(l).hex(),
            "assignment_sha256": sha(r.newc(l)),
            "partition_sha256": sha(node, comm),
        })
    final = r.final_comm()
    q = r.final_modularity()
    doc = {
        "graph": "rmat26_ef16_float_seed2",
        "generator_sha256": gdig,
        "n": int(n), "e": int(len(src)), "scale": r.scale,
        "levels": levels,
        "final_sha256": sha(final),
        "communities": int(final.max()) + 1,
        "final_modularity": q.hex(),
        "oracle_threads": threads,
        "oracle_s": run_s,
    }
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print("[fixture] wrote %s: %d communities, Q %.15f" % (out, doc["communities"], q), flush=True)


if __name__ == "__main__":
    main()
