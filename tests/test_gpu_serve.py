"""Serving (continuous batching): every request's result in the mix - graphs
of different sizes and weight types arriving while others run, admitted into
sets by HBM footprint on two workers - equals the same graph run alone."""
import threading
import time

import numpy as np
import pytest

from conftest import gpu_case
from louvaingpu_amd import Engine, Server, generators
from louvaingpu_amd._lib import LV_ERR_INVALID_INPUT, LV_ERR_NOMEM, LouvainError

pytestmark = pytest.mark.gpu


def _alone(g, gamma):
    with Engine(gamma=gamma) as eng:
        eng.load_edges(*g)
        eng.run()
        return eng.final_assignment(), eng.num_levels(), eng.level_info(eng.num_levels() - 1).q


def _mix():
    params = generators.batch_params(40, seed=11)
    graphs = [(generators.sbm(min(n, 4000), 100, p, p / 50, sd), [1.0, 0.5, 2.0][i % 3])
              for i, (n, p, sd) in enumerate(params)]
    graphs.insert(7, (generators.rmat(14, 16, 0.57, 0.19, 0.19, 5, "float"), 1.0))
    graphs.insert(23, (generators.rmat(13, 16, 0.57, 0.19, 0.19, 6, "int"), 1.0))
    return graphs


def test_results_in_the_mix_equal_alone():
    graphs = _mix()
    gpu_case("serve mix of %d requests, 2 workers" % len(graphs))
    with Server(workers=2, budget_bytes=2 << 30) as srv:
        ids = []
        for i, (g, gamma) in enumerate(graphs):
            ids.append(srv.submit(*g, gamma=gamma))
            if i % 5 == 4:
                time.sleep(0.002)  # later requests arrive while sets run
        results = [srv.wait(r) for r in ids]
        st = srv.stats()
        assert st.completed == len(graphs)
        assert st.sets >= 2
    for (g, gamma), (comm, res) in zip(graphs, results):
        assert res.status == 0
        exp, levels, q = _alone(g, gamma)
        assert np.array_equal(comm, exp)
        assert res.levels == levels
        assert res.q == q
        assert res.communities == int(exp.max()) + 1
        assert res.t_submit_ms <= res.t_start_ms <= res.t_done_ms


def test_concurrent_submitters():
    graphs = _mix()[:24]
    gpu_case("serve concurrent submitters")
    out = {}
    with Server(workers=3, budget_bytes=1 << 30) as srv:
        def client(k):
            for i in range(k, len(graphs), 4):
                g, gamma = graphs[i]
                out[i] = srv.wait(srv.submit(*g, gamma=gamma))[0]
        ths = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    for i, (g, gamma) in enumerate(graphs):
        assert np.array_equal(out[i], _alone(g, gamma)[0]), i


def test_bad_requests_get_their_own_status():
    """An invalid request (an endpoint outside 0..n-1, a negative weight) in a
    set with valid ones fails alone; a request larger than the budget is
    refused at submit; every valid request equals its run alone."""
    params = generators.batch_params(12, seed=21)
    good = [generators.sbm(min(n, 3000), 100, p, p / 50, sd) for n, p, sd in params]
    n0, s0, d0, w0 = good[0]
    bad_id = (n0, s0, np.where(np.arange(len(d0)) == 5, n0, d0).astype(np.int32), w0)
    bad_w = (n0, s0, d0, np.where(np.arange(len(w0)) == 3, -1, w0).astype(np.int32))
    blocker = generators.rmat(16, 16, 0.57, 0.19, 0.19, 9, "int")
    budget = 512 << 20
    huge_n = budget // 320 + 1
    gpu_case("serve bad requests in a mix, 1 worker, budget %d MiB" % (budget >> 20))
    with Server(workers=1, budget_bytes=budget) as srv:
        first = srv.submit(*blocker)  # keeps the worker busy while the mix queues
        ids, kinds = [], []
        for i, g in enumerate(good):
            ids.append(srv.submit(*g))
            kinds.append(("good", i))
            if i == 3:
                ids.append(srv.submit(*bad_id))
                kinds.append(("bad", None))
            if i == 7:
                ids.append(srv.submit(*bad_w))
                kinds.append(("bad", None))
        with pytest.raises(LouvainError) as ex:
            srv.submit(huge_n, np.zeros(0, np.int32), np.zeros(0, np.int32),
                       np.zeros(0, np.int32))
        assert ex.value.status == LV_ERR_NOMEM
        srv.wait(first)
        mixed = 0
        for rid, (kind, i) in zip(ids, kinds):
            if kind == "bad":
                with pytest.raises(LouvainError) as ex:
                    srv.wait(rid)
                assert ex.value.status == LV_ERR_INVALID_INPUT
                continue
            comm, res = srv.wait(rid)
            assert res.status == 0
            mixed += res.batch_size > 1
            assert np.array_equal(comm, _alone(good[i], 1.0)[0]), i
        assert mixed > 0  # some valid requests ran in a set with a bad one


def test_poisson_stream_of_the_serving_workload():
    """Requests drawn as in the serving workload (SBMs with n log-uniform in
    [100, 50000], seed 4, plus R-MAT scale-20 float graphs) arriving as a
    Poisson stream at 200 requests/s on two workers, with a burst of four
    requests at once every 16 (so that sets of several requests form however
    fast the workers are): every result equals the request run alone."""
    rng = np.random.Generator(np.random.PCG64(4))
    nreq = 80
    gaps = rng.exponential(1.0 / 200.0, nreq)
    gaps[np.arange(nreq) % 16 >= 13] = 0.0  # requests 13-15 of every 16 follow 12 at once
    params = generators.batch_params(nreq, seed=4)
    rmat = [generators.rmat(20, 16, 0.57, 0.19, 0.19, sd, "float") for sd in (4, 5)]
    reqs = [generators.sbm(n, 100, p, p / 50, sd) for n, p, sd in params]
    reqs[11], reqs[47] = rmat  # two R-MAT 20 requests inside the stream
    gpu_case("serve poisson stream: %d requests, 2 R-MAT 20" % nreq)
    with Server(workers=2, budget_bytes=8 << 30) as srv:
        ids = []
        for i, g in enumerate(reqs):
            time.sleep(gaps[i])
            ids.append(srv.submit(*g))
        results = [srv.wait(r) for r in ids]
    sizes = []
    for g, (comm, res) in zip(reqs, results):
        assert res.status == 0
        assert np.array_equal(comm, _alone(g, 1.0)[0])
        sizes.append(res.batch_size)
    assert max(sizes) > 1  # requests did share sets
