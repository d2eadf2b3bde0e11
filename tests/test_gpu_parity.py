"""Parity of the HIP engine (through its C ABI) with the oracle.

Integer/index results are compared bit for bit: every pass's next[] of every
level, the Q each break test compares (FP64, bit-equal: both sides evaluate
the same expression on the same exact integers), the level's returned Q,
renumbered assignments, Partition arrays, coarsened CSR rows and the final
flattened partition.  The final Newman modularity, which the engine forms
from exact int64 aggregates and the oracle sums naively in FP64, must agree
within 1e-9 relative.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import gpu_case
from louvaingpu_amd import Engine, generate_dendogram_phases, generators
from louvaingpu_amd._lib import LouvainError, LV_ERR_INVALID_INPUT, LV_ERR_RANGE
from test_oracle_pins import _random_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
Q_REL = 1e-9


def merged_rows(n, rp, col, w):
    """(row, col, sum w) over non-self entries, sorted."""
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    col = col.astype(np.int64)
    keep = rows != col
    key = rows[keep] * n + col[keep]
    ww = w[keep]
    order = np.argsort(key, kind="stable")
    key, ww = key[order], ww[order]
    if len(key) == 0:
        return key, ww
    heads = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    return key[heads], np.add.reduceat(ww, heads)


def assert_run_matches(eng, r, n, trace=True):
    L = eng.num_levels()
    assert L == r.num_levels
    for l in range(L):
        info = eng.level_info(l)
        assert info.n == r.level_n(l), l
        assert info.k == r.level_k(l), l
        assert info.m_units == r.level_m(l), l
        assert info.passes == r.level_passes(l), l
        for p in range(1, info.passes + 1):
            assert eng.trace_pass_q(l, p) == r.pass_q(l, p), (l, p)
            if trace:
                got = eng.trace_next(l, p)
                exp = r.pass_next(l, p)
                if not np.array_equal(got, exp):
                    bad = np.flatnonzero(got != exp)
                    raise AssertionError("level %d pass %d: %d of %d differ, first %s: %s vs %s"
                                         % (l, p, len(bad), len(exp), bad[:5], got[bad[:5]],
                                            exp[bad[:5]]))
        assert info.q == r.level_q(l), l
        assert info.lower == r.level_lower(l), l
        assert np.array_equal(eng.level_assignment(l), r.newc(l)), l
        node, comm = eng.level_partition(l)
        onode, ocomm = r.partition(l)
        assert np.array_equal(node, onode), l
        assert np.array_equal(comm, ocomm), l
    assert np.array_equal(eng.final_assignment(), r.final_comm())
    qe = eng.final_modularity()
    qo = r.final_modularity()
    assert math.isclose(qe, qo, rel_tol=Q_REL, abs_tol=Q_REL), (qe, qo)


def run_both(n, src, dst, w, ids=None, gamma=1.0, threads=8, **tunables):
    r = oracle.run(n, src, dst, w, ids=ids, gamma=gamma, trace=True, threads=threads)
    eng = Engine(gamma=gamma, trace=True, **tunables)
    eng.load_edges(n, src, dst, w, ids=ids)
    eng.run()
    return eng, r


def test_appendix_b_known_answer_on_gpu():
    gpu_case("appendix_b")
    with open(os.path.join(GOLDEN, "appendix_b.json")) as f:
        g = json.load(f)
    w = np.array(g["w"], np.int32)
    eng, r = run_both(g["n"], g["src"], g["dst"], w)
    with eng:
        assert_run_matches(eng, r, g["n"])
        for l, exp in enumerate(g["levels"]):
            for p, ep in enumerate(exp["passes"], start=1):
                assert eng.trace_next(l, p).tolist() == ep["next"]
            node, comm = eng.level_partition(l)
            assert node.tolist() == exp["partition_node"]
            assert comm.tolist() == exp["partition_comm"]
        node, comm = eng.flatten(1)
        assert comm.tolist() == g["flatten_level1"]
        assert math.isclose(eng.final_modularity(), 5 / 14, rel_tol=1e-12)


@pytest.mark.parametrize("seed", range(0, 300, 3))
def test_random_small_graphs(seed):
    n, src, dst, w, ids, gamma = _random_graph(seed)
    gpu_case("small seed=%d n=%d e=%d gamma=%s ids=%s" % (seed, n, len(src), gamma, ids is not None))
    esrc = ids[src] if ids is not None else src
    edst = ids[dst] if ids is not None else dst
    r = oracle.run(n, esrc, edst, w, ids=ids, gamma=gamma, trace=True)
    with Engine(gamma=gamma, trace=True) as eng:
        eng.load_edges(n, esrc, edst, w, ids=ids)
        eng.run()
        if len(src) == 0:
            assert eng.num_levels() == 1
            assert eng.final_assignment().tolist() == list(range(n))
            return
        assert_run_matches(eng, r, n)


def test_sbm_10k_every_pass_and_csr():
    gpu_case("sbm_10k phases")
    n, src, dst, w = generators.sbm_10k()
    r = oracle.run(n, src, dst, w, trace=True)
    with Engine(trace=True) as eng:
        eng.load_edges(n, src, dst, w)
        # phase by phase, checking the CSR the engine builds at each level
        eng.CalculateNeighbour()
        level = 0
        mod = None
        while True:
            _, ln, nnz = eng.current_level()
            rp, col, cw = eng.level_csr()
            orp, ocol, ocw = r.csr(level)
            gk, gw = merged_rows(ln, rp, col, cw)
            ok, ow = merged_rows(ln, orp, ocol, ocw)
            assert np.array_equal(gk, ok) and np.array_equal(gw, ow), level
            eng.initCUDA()
            k, loop, selfsum = eng.level_vertex()
            ok_, oloop = r.vertex(level)
            assert np.array_equal(k, ok_) and np.array_equal(loop, oloop), level
            if mod is None:
                mod = eng.modularityCUDA()
                assert mod == r.level_q_init(0)
            q, passes = eng.oneLevelCUDA(mod)
            assert q == r.level_q(level)
            eng.renumberCUDA()
            if level > 0 and q - mod < 1e-7:
                break
            mod = q
            eng.inducedGraphCUDA()
            eng.CalculateNeighbour()
            level += 1
        assert_run_matches(eng, r, n)


@pytest.mark.parametrize("scale,weights,seed", [(12, "int", 1), (14, "float", 2), (16, "int", 3)])
def test_rmat_with_hubs(scale, weights, seed):
    gpu_case("rmat scale=%d %s" % (scale, weights))
    n, src, dst, w = generators.rmat(scale, 16, 0.57, 0.19, 0.19, seed, weights)
    eng, r = run_both(n, src, dst, w)
    with eng:
        assert_run_matches(eng, r, n)


def test_resolution_sweep_on_sbm():
    n, src, dst, w = generators.sbm(3000, 100, 0.08, 0.0016, 11)
    for gamma in generators.resolutions(16)[::3]:
        gpu_case("sbm3000 gamma=%g" % gamma)
        eng, r = run_both(n, src, dst, w, gamma=gamma)
        with eng:
            assert_run_matches(eng, r, n)


def test_arbitrary_ids_float_weights_and_loops():
    gpu_case("ids+float+loops")
    rng = np.random.Generator(np.random.PCG64(99))
    n = 500
    ids = (rng.permutation(n) * 7 - 1000).astype(np.int32)
    e = 3000
    s = rng.integers(0, n, e)
    t = rng.integers(0, n, e)
    t[:50] = s[:50]                                  # self-loops, some repeated
    w = (rng.integers(0, 40, e) * 0.125).astype(np.float32)  # zeros included
    eng, r = run_both(n, ids[s], ids[t], w, ids=ids)
    with eng:
        assert_run_matches(eng, r, n)
        node, comm = eng.flatten(eng.num_levels() - 1)
        onode, ocomm = r.flatten(r.num_levels - 1)
        assert np.array_equal(node, onode) and np.array_equal(comm, ocomm)


@pytest.mark.parametrize("gamma", [0.5, 1.0, 2.0])
def test_unsorted_list_with_repeated_head(gamma):
    """One wavefront sees both order faults of the edge list: a repeated pair
    at the head (same source, destination not increasing) and a later edge
    whose source decreases.  Both must reach the flags so the list takes the
    full row sort."""
    gpu_case("unsorted list with repeated head, gamma %g" % gamma)
    src = np.array([3, 3, 0, 0, 3, 3, 0, 1], np.int32)
    dst = np.array([3, 1, 1, 3, 1, 1, 2, 2], np.int32)
    w = np.array([1, 1, 2, 2, 2, 2, 2, 1], np.int32)
    eng, r = run_both(4, src, dst, w, gamma=gamma)
    with eng:
        assert_run_matches(eng, r, 4)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
@pytest.mark.parametrize("wmax_log2", [18, 11])
def test_wide_weight_units(order, wmax_log2):
    """Weights up to 2^18 next to 2^-20 make the weight unit 2^-20, so the
    total degree is about 2^57 units: with n = 2^15 the per-community size
    delta no longer fits beside the degree delta in one 64-bit word and the
    sweep keeps them in separate arrays; the level is not narrow, and the
    CSR entries (weights up to 2^38 units) take the two-array row sort of
    either edge order instead of one word per entry and keep 8-byte weights.
    With weights up to 2^11 every level-0 weight fits 32 bits (2^31 units):
    level 0 keeps 4-byte row weights while the coarser levels, whose
    weights the total degree (about 2^50) no longer bounds by 2^32, keep
    8-byte ones."""
    gpu_case("wide weight units, %s, weights below 2^%d" % (order, wmax_log2))
    n, src, dst, _ = generators.rmat(15, 16, 0.57, 0.19, 0.19, 11, "int")
    rng = np.random.Generator(np.random.PCG64(5))
    w = rng.integers(1, 1 << wmax_log2, len(src)).astype(np.float32)
    w[rng.integers(0, len(src), 64)] = np.float32(2.0 ** -20)
    if order == "shuffled":
        p = rng.permutation(len(src))
        src, dst, w = src[p], dst[p], w[p]
    eng, r = run_both(n, src, dst, w)
    with eng:
        k = np.asarray(r.vertex(0)[0], np.float64)
        assert math.log2(k.sum()) + (n.bit_length() + 1) > 62  # the unpacked case
        assert_run_matches(eng, r, n)


def test_repeat_runs_are_identical():
    gpu_case("repeat runs")
    n, src, dst, w = generators.rmat(14, 16, 0.57, 0.19, 0.19, 7, "int")
    with Engine() as eng:
        eng.load_edges(n, src, dst, w)
        eng.run()
        a = eng.final_assignment()
        qa = eng.final_modularity()
        for _ in range(2):
            eng.run()
            assert np.array_equal(eng.final_assignment(), a)
            assert eng.final_modularity() == qa


def test_invalid_input_is_reported():
    gpu_case("invalid inputs")
    with Engine() as eng:
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(3, [0, 1], [1, 3], np.ones(2, np.int32))
        assert ex.value.status == LV_ERR_INVALID_INPUT
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(3, [0, 1], [1, 2], np.array([1, -2], np.int32))
        assert ex.value.status == LV_ERR_INVALID_INPUT
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(3, [0, 1], [1, 2], np.array([1, np.inf], np.float32))
        assert ex.value.status == LV_ERR_INVALID_INPUT
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(3, [5, 6], [6, 7], np.ones(2, np.int32), ids=[5, 6, 6])
        assert ex.value.status == LV_ERR_INVALID_INPUT
        with pytest.raises(LouvainError) as ex:
            eng.load_edges(3, [0, 1], [1, 2], np.array([1e30, 1e-30], np.float32))
        assert ex.value.status == LV_ERR_RANGE
        # a valid load afterwards works, and an edgeless graph is the identity
        eng.load_edges(4, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
        assert eng.run() == 1
        assert eng.final_assignment().tolist() == [0, 1, 2, 3]


def test_phase_api_matches_run():
    gpu_case("phase api vs run")
    n, src, dst, w = generators.sbm(2000, 100, 0.1, 0.002, 5)
    r = oracle.run(n, src, dst, w)
    with Engine() as eng:
        eng.load_edges(n, src, dst, w)
        dendro = generate_dendogram_phases(eng)
        assert len(dendro) == r.num_levels
        for l, (node, comm) in enumerate(dendro):
            onode, ocomm = r.partition(l)
            assert np.array_equal(node, onode) and np.array_equal(comm, ocomm)


@pytest.mark.timeout(600)
def test_rmat22_full_size():
    """The R-MAT scale-22 benchmark graph end to end: every sweep's next[] of
    every level, every pass Q bit-equal, Partition arrays and final partition
    identical, modularity within 1e-9."""
    gpu_case("rmat22 full")
    n, src, dst, w = generators.rmat22()
    r = oracle.run(n, src, dst, w, trace=True, threads=len(os.sched_getaffinity(0)))
    with Engine(trace=True) as eng:
        eng.load_edges(n, src, dst, w)
        eng.run()
        assert_run_matches(eng, r, n, trace=True)


def _hub_graph(seed, hubs=3, hub_deg=9000, n=30000):
    rng = np.random.Generator(np.random.PCG64(seed))
    src, dst = [], []
    for h in range(hubs):
        nb = rng.choice(np.arange(hubs, n), hub_deg, replace=False)
        src += [h] * hub_deg
        dst += nb.tolist()
    e = 3 * n
    src += rng.integers(hubs, n, e).tolist()
    dst += rng.integers(hubs, n, e).tolist()
    src += [0, 1]            # self-loops on two hubs: the virtual own-community entry
    dst += [0, 1]
    w = rng.integers(1, 5, len(src)).astype(np.int32)
    return n, np.array(src, np.int32), np.array(dst, np.int32), w


def test_hub_partitions():
    gpu_case("hub partitions")
    n, src, dst, w = _hub_graph(5)
    eng, r = run_both(n, src, dst, w)
    with eng:
        assert_run_matches(eng, r, n)


def test_star_with_long_chunks():
    """A hub with 1.2 M neighbours (P = 586 hash partitions) and the largest
    community of level 0 (2.25 M entries, P = 1098): both partition pipelines
    cut chunks of 8 entries per partition, longer than the 4096-entry minimum."""
    gpu_case("star with long chunks")
    rng = np.random.Generator(np.random.PCG64(17))
    leaves = 1_200_000
    n = leaves + 1
    src = np.concatenate([np.zeros(leaves, np.int64), rng.integers(1, n, 100_000)])
    dst = np.concatenate([np.arange(1, n), rng.integers(1, n, 100_000)])
    w = rng.integers(1, 4, len(src)).astype(np.int32)
    eng, r = run_both(n, src.astype(np.int32), dst.astype(np.int32), w)
    with eng:
        assert_run_matches(eng, r, n, trace=False)


@pytest.mark.parametrize("hub_q,part,chunk", [(512, 1024, 8), (64, 256, 1), (1024, 4096, 64)])
def test_launch_shapes_do_not_change_results(hub_q, part, chunk):
    """Any hub threshold, partition target and chunk length (the autotuner's
    knobs) gives the oracle's results: more rows and coarsening segments take
    the partition pipelines at low thresholds."""
    gpu_case("launch shapes hub_q=%d part=%d chunk=%d" % (hub_q, part, chunk))
    n, src, dst, w = generators.rmat(15, 16, 0.57, 0.19, 0.19, 13, "int")
    eng, r = run_both(n, src, dst, w, hub_q=hub_q, hub_part_target=part, chunk_mul=chunk)
    with eng:
        assert_run_matches(eng, r, n)


@pytest.mark.parametrize("wtype", ["int", "float"])
def test_rows_of_2049_to_4096_in_block_tables(wtype):
    """hub_q = 4096: rows of 2049..4096 candidates take one 512-lane block
    with an 8192-slot LDS table instead of the hub partitions (4-byte table
    sums on the int graph, 8-byte ones on the float graph, whose total degree
    in units exceeds 2^32); coarsening segments above 2048 entries keep the
    partitions."""
    gpu_case("block tables for rows of 2049..4096, %s weights" % wtype)
    n, src, dst, w = generators.rmat(17, 16, 0.57, 0.19, 0.19, 21, wtype)
    ns = src != dst
    q = np.bincount(src[ns], minlength=n) + np.bincount(dst[ns], minlength=n)
    assert ((q > 2048) & (q <= 4096)).sum() > 10
    eng, r = run_both(n, src, dst, w, hub_q=4096, timing=True)
    with eng:
        assert eng.stats().tier_launches[10] > 0  # tier 11: rows of 2049..4096
        assert_run_matches(eng, r, n)


def test_hub_global_fallback():
    """One partition per hub holds ~9000 distinct communities, more than the
    4096-slot LDS table: the partition is redone in its global table."""
    gpu_case("hub global fallback")
    n, src, dst, w = _hub_graph(6)
    eng, r = run_both(n, src, dst, w, hub_part_target=1000000)
    with eng:
        assert_run_matches(eng, r, n)


def test_coarsening_global_fallback():
    """A hub tied to 6000 heavy pairs: each pair becomes a community, so the
    hub's community coarsens into one partition (partition target 10^6) with
    6000 distinct neighbour communities, more than its 4096-slot LDS table:
    the partition is aggregated in its global table and written into the row."""
    gpu_case("coarsening global fallback")
    k = 6000
    u = 1 + 2 * np.arange(k, dtype=np.int32)
    src = np.concatenate([u, np.zeros(k, np.int32)])
    dst = np.concatenate([u + 1, u])
    w = np.concatenate([np.full(k, 100, np.int32), np.ones(k, np.int32)])
    eng, r = run_both(1 + 2 * k, src, dst, w, hub_part_target=1000000)
    with eng:
        assert eng.level_info(1).n > 4096
        assert_run_matches(eng, r, 1 + 2 * k)


@pytest.mark.parametrize("variant", ["shuffled", "reversed_pairs", "dups_loops_sorted"])
def test_csr_paths_agree(variant):
    """The CSR of an edge list sorted by source (forward entries kept in place,
    only reverse entries row-sorted) and of any other order (full row sort)
    give the oracle's results; rows with repeated neighbours take the
    aggregating first pass."""
    n, src, dst, w = generators.rmat(13, 16, 0.57, 0.19, 0.19, 21, "int")
    rng = np.random.Generator(np.random.PCG64(5))
    if variant == "shuffled":
        p = rng.permutation(len(src))
        src, dst, w = src[p], dst[p], w[p]
    elif variant == "reversed_pairs":
        # sorted by source but not canonical: some edges stored as (max, min)
        flip = rng.uniform(size=len(src)) < 0.5
        s2 = np.where(flip, dst, src)
        d2 = np.where(flip, src, dst)
        o = np.lexsort((d2, s2))
        src, dst, w = s2[o], d2[o], w[o]
    else:
        # duplicate edges and extra self-loops, still sorted by source
        k = len(src) // 10
        pick = rng.integers(0, len(src), k)
        loops = rng.integers(0, n, k // 4).astype(np.int32)
        src = np.concatenate([src, src[pick], loops])
        dst = np.concatenate([dst, dst[pick], loops])
        w = np.concatenate([w, w[pick], np.ones(len(loops), np.int32)])
        o = np.argsort(src, kind="stable")
        src, dst, w = src[o], dst[o], w[o]
    gpu_case("csr paths %s" % variant)
    eng, r = run_both(n, src.astype(np.int32), dst.astype(np.int32), w.astype(np.int32))
    try:
        assert_run_matches(eng, r, n)
    finally:
        eng.close()


def test_run_partitions_delivers_the_dendrogram():
    """lv_run_partitions copies every level's Partition and the final partition
    to the host during the run: equal to lv_level_partition / lv_final_assignment
    after lv_run, and to the oracle; fewer buffers than levels leave the rest
    on the device; an edgeless graph gives its one identity level."""
    from louvaingpu_amd.engine import pinned_empty
    n, src, dst, w = generators.rmat(14, 16, 0.57, 0.19, 0.19, 9, "int")
    gpu_case("run_partitions rmat14 n=%d e=%d" % (n, len(src)))
    r = oracle.run(n, src, dst, w)
    with Engine() as eng:
        eng.load_edges(n, src, dst, w)
        L = r.num_levels
        nodes = [pinned_empty(n, np.int32) for _ in range(L)]
        comms = [np.full(n, -7, np.int32) for _ in range(L)]  # pageable works too
        fin = pinned_empty(n, np.int32)
        assert eng.run_partitions(nodes, comms, fin) == L
        for l in range(L):
            k = r.level_n(l)
            onode, ocomm = r.partition(l)
            assert np.array_equal(nodes[l][:k], onode) and np.array_equal(comms[l][:k], ocomm), l
        assert np.array_equal(fin, r.final_comm())
        # one buffer pair: level 0 delivered, the others fetched afterwards
        fin2 = np.zeros(n, np.int32)
        n0, c0 = np.zeros(n, np.int32), np.zeros(n, np.int32)
        assert eng.run_partitions([n0], [c0], fin2) == L
        assert np.array_equal(n0, r.partition(0)[0]) and np.array_equal(c0, r.partition(0)[1])
        assert np.array_equal(eng.level_partition(L - 1)[1], r.partition(L - 1)[1])
        assert np.array_equal(fin2, fin)
        eng.load_edges(6, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
        a, b, f = np.zeros(6, np.int32), np.zeros(6, np.int32), np.zeros(6, np.int32)
        assert eng.run_partitions([a], [b], f) == 1
        assert f.tolist() == list(range(6)) and b.tolist() == list(range(6))


@pytest.mark.parametrize("wmax", [(1 << 16) - 1, 1 << 16])
def test_level0_entry_layout_at_the_32_bit_boundary(wmax):
    """A single graph's level 0 built from a source-sorted list keeps each row
    entry as one word col << wb | weight when bits(n - 1) + bits(largest
    weight) <= 32 (the partition entries and the coarsening's member-order
    entries of level 0 likewise), else columns and weights apart: n = 2^16
    with weights up to 2^16 - 1 sits exactly on the packed side (16 + 16 bits),
    2^16 one bit past it.  Both give the oracle's CSR, passes and levels."""
    gpu_case("level-0 entry layout, n=2^16, largest weight %d" % wmax)
    rng = np.random.Generator(np.random.PCG64(29))
    n = 1 << 16
    # hubs (rows past the block tiers) over a sparse random graph, canonical
    # (src < dst) and sorted by (src, dst) so the sorted CSR build runs
    a = np.concatenate([np.repeat(np.arange(8), 6000), rng.integers(0, n, 400_000)])
    b = np.concatenate([rng.integers(8, n, 48_000), rng.integers(0, n, 400_000)])
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keep = lo != hi
    key = np.unique(lo[keep].astype(np.int64) * n + hi[keep])
    src = (key // n).astype(np.int32)
    dst = (key % n).astype(np.int32)
    w = rng.integers(1, 50, len(src)).astype(np.int32)
    w[rng.integers(0, len(w), 64)] = wmax  # the largest weight sets the layout
    r = oracle.run(n, src, dst, w, trace=True)
    with Engine(trace=True) as eng:
        eng.load_edges(n, src, dst, w)
        eng.CalculateNeighbour()
        rp, col, cw = eng.level_csr()
        orp, ocol, ocw = r.csr(0)
        gk, gw = merged_rows(n, rp, col, cw)
        ok, ow = merged_rows(n, orp, ocol, ocw)
        assert np.array_equal(gk, ok) and np.array_equal(gw, ow)
    eng, r = run_both(n, src, dst, w)
    with eng:
        assert_run_matches(eng, r, n)
