"""Graph sets (lv_batch_*): every component of a batched run must equal the
single-graph run of its graph at its resolution - levels, per-level k, passes,
Q and lower bit for bit, every level's renumbered assignment and the final
partition - and, on the SBM workload, the oracle."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from louvaingpu_amd import Engine, GraphSet, generators
from louvaingpu_amd._lib import LouvainError, LV_ERR_INVALID_INPUT
from test_oracle_pins import _random_graph

pytestmark = pytest.mark.gpu


def _single(g, gamma, **cfg):
    n, src, dst, w = g
    eng = Engine(gamma=gamma, **cfg)
    eng.load_edges(n, src, dst, w)
    eng.run()
    return eng


def assert_component_matches(gs, comp, eng):
    L = eng.num_levels()
    assert gs.num_levels(comp) == L, comp
    for l in range(L):
        a, b = gs.level_info(comp, l), eng.level_info(l)
        assert (a.n, a.k, a.passes, a.m_units, a.scale) == (b.n, b.k, b.passes, b.m_units, b.scale), (comp, l)
        assert a.q == b.q and a.lower == b.lower, (comp, l, a.q, b.q, a.lower, b.lower)
        assert np.array_equal(gs.level_assignment(comp, l), eng.level_assignment(l)), (comp, l)
    assert np.array_equal(gs.final(comp), eng.final_assignment()), comp


def _local_graph(seed):
    n, src, dst, w, ids, gamma = _random_graph(seed)
    return (n, src.astype(np.int32), dst.astype(np.int32), w)


@pytest.mark.parametrize("kind", ["int", "float"])
def test_random_small_sets_match_single_runs(kind):
    # tiny graphs with zero weights, self-loops, duplicates, edgeless members
    seeds = [s for s in range(200) if (s % 4 in (0, 1)) == (kind == "int")][:40]
    graphs = [_local_graph(s) for s in seeds]
    gammas = [0.5, 1.0, 2.0]
    gpu_case("batch random %s: %d graphs x %d gammas" % (kind, len(graphs), len(gammas)))
    with GraphSet() as gs:
        gs.load(graphs, gammas)
        gs.run()
        for gi, g in enumerate(graphs):
            for ri, gamma in enumerate(gammas):
                eng = _single(g, gamma)
                try:
                    assert_component_matches(gs, gi * len(gammas) + ri, eng)
                finally:
                    eng.close()


def test_sbm_set_matches_single_runs_and_oracle():
    params = generators.batch_params(24, seed=3)
    graphs = [generators.sbm(min(n, 3000), 100, p, p / 50, sd) for n, p, sd in params]
    gammas = generators.resolutions(16)[::5]  # 0.25, 0.63, 1.59, 4.0
    gpu_case("batch sbm: %d graphs x %d gammas" % (len(graphs), len(gammas)))
    with GraphSet() as gs:
        gs.load(graphs, gammas)
        gs.run()
        for gi, g in enumerate(graphs):
            for ri, gamma in enumerate(gammas):
                comp = gi * len(gammas) + ri
                eng = _single(g, gamma)
                try:
                    assert_component_matches(gs, comp, eng)
                finally:
                    eng.close()
                if gi < 4:
                    r = oracle.run(*g, gamma=gamma, threads=8)
                    assert np.array_equal(gs.final(comp), r.final_comm()), comp
                    assert gs.num_levels(comp) == r.num_levels


def test_set_with_hub_rows_matches_single_runs():
    # R-MAT members exercise the block and hub sweep tiers inside a set
    graphs = [generators.rmat(13, 16, 0.57, 0.19, 0.19, s, "int") for s in range(3)]
    graphs.append(generators.rmat(12, 16, 0.57, 0.19, 0.19, 7, "int"))
    gammas = [1.0, 1.7]
    gpu_case("batch rmat: %d graphs x %d gammas" % (len(graphs), len(gammas)))
    with GraphSet() as gs:
        gs.load(graphs, gammas)
        gs.run()
        for gi, g in enumerate(graphs):
            for ri, gamma in enumerate(gammas):
                eng = _single(g, gamma)
                try:
                    assert_component_matches(gs, gi * len(gammas) + ri, eng)
                finally:
                    eng.close()


def test_set_limits_apply_per_component():
    graphs = [generators.sbm(1000, 100, 0.08, 0.0016, s) for s in range(5)]
    gpu_case("batch limits")
    for cfg in ({"max_levels": 1}, {"max_passes": 1}, {"min_gain": 1e-3}):
        with GraphSet(**cfg) as gs:
            gs.load(graphs, [1.0, 3.0])
            gs.run()
            for gi, g in enumerate(graphs):
                for ri, gamma in enumerate([1.0, 3.0]):
                    eng = _single(g, gamma, **cfg)
                    try:
                        assert_component_matches(gs, gi * 2 + ri, eng)
                    finally:
                        eng.close()


def test_set_final_all_and_repeat_runs():
    graphs = [generators.sbm(500 + 100 * i, 100, 0.06, 0.0012, i) for i in range(6)]
    gpu_case("batch final_all / repeat")
    with GraphSet() as gs:
        gs.load(graphs, [0.8, 1.2])
        gs.run()
        first = gs.final_all()
        parts = np.concatenate([gs.final(c) for c in range(gs.ncomp)])
        assert np.array_equal(first, parts)
        gs.run()
        assert np.array_equal(gs.final_all(), first)


def test_set_edgeless_and_invalid():
    gpu_case("batch edgeless / invalid")
    empty = (5, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    with GraphSet() as gs:
        gs.load([empty, empty], [1.0])
        gs.run()
        assert gs.num_levels(0) == 1
        assert gs.final(1).tolist() == list(range(5))
    bad = (3, np.array([0, 3], np.int32), np.array([1, 1], np.int32), np.ones(2, np.int32))
    with GraphSet() as gs:
        with pytest.raises(LouvainError) as ei:
            gs.load([bad], [1.0])
        assert ei.value.status == LV_ERR_INVALID_INPUT


def test_reloading_one_context_with_growing_sets():
    # later loads need more components and more staged edges than earlier ones,
    # so the pinned buffers of the context grow between runs
    sizes = [2, 30, 200]
    params = generators.batch_params(max(sizes), seed=5)
    graphs = [generators.sbm(min(n, 1500), 100, p, p / 50, sd) for n, p, sd in params]
    gpu_case("batch reload growing sets %s" % sizes)
    with GraphSet() as gs:
        for k in sizes:
            gs.load(graphs[:k], [1.0, 2.0])
            gs.run()
            for gi in (0, k // 2, k - 1):
                for ri, gamma in enumerate([1.0, 2.0]):
                    eng = _single(graphs[gi], gamma)
                    try:
                        assert np.array_equal(gs.final(gi * 2 + ri), eng.final_assignment()), (k, gi)
                    finally:
                        eng.close()


def test_batch_workload_sample_full_size():
    """Members of the batching workload itself (4,096 SBMs with n log-uniform in
    [100, 50000], seed 3) at full size and all 16 resolutions: every 256th
    graph, each component equal to its single-graph run, the largest two also
    to the oracle at three resolutions."""
    params = generators.batch_params(4096, seed=3)[::256]
    graphs = [generators.sbm(n, 100, p, p / 50, sd) for n, p, sd in params]
    gammas = generators.resolutions(16)
    gpu_case("batch workload sample: %d graphs (n %d..%d) x %d gammas"
             % (len(graphs), min(g[0] for g in graphs), max(g[0] for g in graphs), len(gammas)))
    with GraphSet() as gs:
        gs.load(graphs, gammas)
        gs.run()
        for gi, g in enumerate(graphs):
            for ri, gamma in enumerate(gammas):
                eng = _single(g, gamma)
                try:
                    assert_component_matches(gs, gi * len(gammas) + ri, eng)
                finally:
                    eng.close()
        big = sorted(range(len(graphs)), key=lambda i: -graphs[i][0])[:2]
        for gi in big:
            for ri in (0, 7, 15):
                r = oracle.run(*graphs[gi], gamma=gammas[ri], threads=8)
                assert np.array_equal(gs.final(gi * len(gammas) + ri), r.final_comm()), (gi, ri)
