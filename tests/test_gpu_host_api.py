"""The C++ host layer (Graph / generate_dendogram / partitionAtLevel /
getPartitions / countPartitions) and the start() dict API, on the GPU,
against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from louvaingpu_amd import Engine, count_partitions, generate_dendogram, generators, start
from louvaingpu_amd._lib import LV_SWEEP_TIERS

pytestmark = pytest.mark.gpu


def test_generate_dendogram_matches_oracle():
    gpu_case("host generate_dendogram sbm_10k")
    n, src, dst, w = generators.sbm_10k()
    r = oracle.run(n, src, dst, w)
    d = generate_dendogram(n, src, dst, w.astype(np.float32))
    try:
        assert len(d) == r.num_levels
        for l in range(len(d)):
            node, comm = d.levels[l]
            onode, ocomm = r.partition(l)
            assert np.array_equal(node, onode) and np.array_equal(comm, ocomm)
        for l in range(len(d)):
            node, comm = d.partition_at_level(l)
            onode, ocomm = r.flatten(l)
            assert np.array_equal(node, onode) and np.array_equal(comm, ocomm)
        node, comm = d.get_partitions()        # reference: level `0 || len-1` -> 1
        onode, ocomm = r.flatten(1)
        assert np.array_equal(comm, ocomm)
        node, comm = d.final_partition()
        onode, ocomm = r.flatten(r.num_levels - 1)
        assert np.array_equal(comm, ocomm)
    finally:
        d.close()


def test_start_dict_api():
    gpu_case("start() dict api")
    rng = np.random.Generator(np.random.PCG64(3))
    n = 300
    ids = (np.arange(n) * 3 + 17).astype(np.int32)
    n_, s, t, w = generators.sbm(n, 30, 0.3, 0.01, 9)
    nodes = [{"id": int(i)} for i in ids]
    edges = [{"source": int(ids[a]), "target": int(ids[b]), "weight": float(x)}
             for a, b, x in zip(s, t, w)]
    res = start(nodes, edges, [])
    r = oracle.run(n, ids[s], ids[t], w.astype(np.float32), ids=ids)
    onode, ocomm = r.flatten(1)
    assert res["size"] == count_partitions(ocomm)
    for a, c in zip(onode.tolist(), ocomm.tolist()):
        assert res[a] == c
    for nd in nodes:
        assert nd["cluster"] == res[nd["id"]]
    # circle layout of the reference's layoutProcess on the same partition
    from test_layout import reference_layout
    _, _, _, x, y = reference_layout(ids.tolist(), onode.tolist(), ocomm.tolist(),
                                     count_partitions(ocomm))
    assert [nd["x"] for nd in nodes] == [float(v) for v in x]
    assert [nd["y"] for nd in nodes] == [float(v) for v in y]


def test_undersized_arena_reports_nomem_and_recovers():
    """An arena too small for the graph gives LV_ERR_NOMEM (no abort, no retry
    loop as in LouvainCUDA/LouvainCPP.cpp:921-936); a graph that fits still
    runs on the same context afterwards."""
    from louvaingpu_amd._lib import LV_ERR_NOMEM, LouvainError
    gpu_case("undersized arena")
    big = generators.rmat(14, 16, 0.57, 0.19, 0.19, 3, "int")
    small = generators.sbm(500, 100, 0.1, 0.002, 3)
    with Engine(arena_bytes=4 << 20) as eng:
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(*big)
            eng.run()
        assert ex.value.status == LV_ERR_NOMEM
        eng.load_edges(*small)
        eng.run()
        r = oracle.run(*small)
        assert np.array_equal(eng.final_assignment(), r.final_comm())


def test_timing_stats_count_candidates():
    """With LV_FLAG_TIMING the per-tier stats carry the candidates P of the
    local-moving byte count: on the first pass of a level whose rows list each
    neighbour once, every row entry is a candidate (plus the own community of
    a vertex with a self-loop), and no pass can have more candidates than row
    entries plus vertices."""
    gpu_case("timing stats with candidate counts")
    n, src, dst, w = generators.rmat(14, 16, 0.57, 0.19, 0.19, 3, "int")
    with Engine(timing=True, max_levels=1, max_passes=1) as eng:
        eng.load_edges(n, src, dst, w)
        eng.run()
        st = eng.stats()
        info = eng.level_info(0)
        cand = sum(st.tier_cand[t] for t in range(LV_SWEEP_TIERS))
        nnz = info.nnz
        ns = src != dst
        deg = np.bincount(src[ns], minlength=n) + np.bincount(dst[ns], minlength=n)
        has_loop = np.bincount(src[~ns], minlength=n) > 0
        loops = int(np.count_nonzero(has_loop & (deg > 0)))  # rows without neighbours never sweep
        assert cand == nnz + loops  # one pass, rows unique: every entry, plus own for loops
        for t in range(LV_SWEEP_TIERS):
            if st.tier_launches[t]:
                assert st.tier_bytes[t] >= 8 * st.tier_cand[t]
    with Engine(timing=True) as eng:
        eng.load_edges(n, src, dst, w)
        eng.run()
        st = eng.stats()
        bound = sum(eng.level_info(l).nnz + eng.level_info(l).n for l in range(eng.num_levels()))
        assert 0 < sum(st.tier_cand[t] for t in range(LV_SWEEP_TIERS)) <= 2 * bound
