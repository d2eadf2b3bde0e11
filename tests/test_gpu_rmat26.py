"""R-MAT scale 26 with float32 weights (edge factor 16, seed 2; 1.05 G edges,
2.1 G directed CSR entries) on the engine at full size.  Every per-entry array
is then larger than 4 GiB (byte offsets past 2^32), and the reference's element
counts for its slabs, e.g. (E*2)*5 ints, overflow its int arithmetic
(LouvainCUDA/LouvainCPP.cpp:138-148).

The oracle needs minutes and >100 GB of host memory here, so its results are
committed as digests (tests/golden/rmat26f_oracle.json, written by
tests/golden/make_rmat26f_fixture.py): per level n, k, m, passes, the Q of
every break test and the returned Q (bit-equal), SHA-256 of the renumbered
assignment and of the Partition arrays, and of the final partition.  Beside
them, invariants that need no oracle: each level's vertex count is the
previous level's community count, m is conserved through coarsening (no zero
weights), community ids are dense, and a second run is bit-identical.
"""
import hashlib
import json
import math
import os
import time

import numpy as np
import pytest

from conftest import gpu_case
from louvaingpu_amd import Engine, generators

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat26f_oracle.json")
Q_REL = 1e-9


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.timeout(900)
def test_rmat26_float_full_size():
    with open(FIXTURE) as f:
        fx = json.load(f)
    t0 = time.perf_counter()
    n, src, dst, w = generators.rmat26_float()
    gpu_case("rmat26f generated n=%d e=%d in %.1f s" % (n, len(src), time.perf_counter() - t0))
    assert (n, len(src)) == (fx["n"], fx["e"])
    assert 2 * len(src) * 4 > 2 ** 32 and 10 * len(src) > 2 ** 31  # see the docstring
    assert sha(np.int64(n), src, dst, w) == fx["generator_sha256"]
    with Engine() as eng:
        gpu_case("rmat26f load + run")
        eng.load_edges(n, src, dst, w)
        eng.run()
        levels = fx["levels"]
        assert eng.num_levels() == len(levels)
        ks = []
        for l, lv in enumerate(levels):
            info = eng.level_info(l)
            assert (info.n, info.k, info.m_units, info.passes) == \
                (lv["n"], lv["k"], lv["m_units"], lv["passes"]), l
            assert [eng.trace_pass_q(l, p) for p in range(1, info.passes + 1)] == \
                [float.fromhex(q) for q in lv["pass_q"]], l
            assert info.q == float.fromhex(lv["q"]), l
            assert info.lower == float.fromhex(lv["lower"]), l
            assert sha(eng.level_assignment(l)) == lv["assignment_sha256"], l
            assert sha(*eng.level_partition(l)) == lv["partition_sha256"], l
            if l > 0:
                assert info.n == ks[-1]           # n_{l+1} = K_l
                assert info.m_units == levels[0]["m_units"]  # weight conserved
            ks.append(info.k)
        final = eng.final_assignment()
        assert sha(final) == fx["final_sha256"]
        counts = np.bincount(final)
        assert len(counts) == fx["communities"] and counts.min() > 0  # dense ids
        q = eng.final_modularity()
        qo = float.fromhex(fx["final_modularity"])
        assert math.isclose(q, qo, rel_tol=Q_REL, abs_tol=Q_REL), (q, qo)
        gpu_case("rmat26f second run")
        eng.run()
        assert np.array_equal(eng.final_assignment(), final)
        assert [eng.level_info(l).k for l in range(eng.num_levels())] == ks
        assert eng.final_modularity() == q
