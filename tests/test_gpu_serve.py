"""Serving (continuous batching): every request's result in the mix - graphs
of different sizes and weight types arriving while others run, admitted into
sets by HBM footprint on two workers - equals the same graph run alone."""
import threading
import time

import numpy as np
import pytest

from conftest import gpu_case
from louvaingpu_amd import Engine, Server, generators

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
