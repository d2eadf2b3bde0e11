#!/usr/bin/env python3
"""Louvain benchmark: TEPS of the full multi-level Louvain on one MI355X.

One step = one end-to-end run as the metric defines it: the caller's edge list
in page-locked host memory is uploaded, the reference's whole pipeline runs
(CSR build, init, local-moving sweeps, renumbering, coarsening at every level,
flattening), and every level's Partition (node, community) plus the final
flattened partition are copied back into page-locked host buffers; graph
generation is outside the step.  The graph is R-MAT scale 22 (edge factor 16,
a/b/c/d = .57/.19/.19/.05, duplicates merged into int32 weights, seed 1).
value = TEPS = directed CSR entries of the input (2E) / seconds per step.
The same run with the input already resident in HBM is reported beside it
(device_resident_ms).

Besides the JSON line's required fields it reports:
  roofline     the dominant local-moving sweep tier (one kernel, or the four
               kernels of the hub pipeline): algorithmic bytes 12*nnz + 24*N + 8*P
               per launch (P = the distinct (vertex, community) candidates, whose
               community totals the sweep reads; counted by the kernels) / its
               mean HIP-event duration, against 8 TB/s (frac_p0: the same without
               the 8*P term); traffic = measured HBM bytes per launch from the
               committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
  cpu_baseline the oracle (C restatement of the reference semantics, OpenMP)
               on the same graph with every thread this process may use, timed
               on this host; its partition is also the validation of the GPU
               result.  cpu_baseline_1thread: the oracle on one thread over a
               bounded sample (R-MAT scale 20, same parameters); host_cpu: the
               host's model, nproc, affinity and cgroup CPU quota
  rmat26f      the R-MAT scale-26 float32 graph (seed 2, 2.1 G entries): the same
               end-to-end and device-resident times, phases and roofline, its
               final partition checked against the committed oracle digest
  batched      the batching workload: 4096 SBM graphs (n log-uniform in
               [100, 50000], blocks of 100, p_in ~ U[0.02, 0.1], p_out =
               p_in/50, seed 3) x 16 resolutions 0.25..4.0 as graph sets, in
               graph-runs/s, against the plain per-graph launch loop of the same
               engine on every 64th graph (all resolutions), with identical
               partitions required (--workload batch prints it as its own line)
  serving      Poisson arrivals at 200 requests/s of SBMs drawn as above plus
               2% R-MAT scale-20 float graphs, continuous batching on two
               engine contexts: p50/p99 latency, graphs/s, saturated
               throughput against the per-graph loop, and each result equal
               to the request alone (--workload serve)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Kernels of each sweep tier as named in rocprofv3 summaries: the first pass
# of a level (no aggregation) and the later passes.  The last kernel of each
# list runs once per launch of the tier.  Tiers 5-8 share their first-pass
# kernel, and so do tiers 9-10, so their traffic is not separable.
TIER_KERNELS = {
    1: (["k_sweep1_group<4,", "k_sweep_group<4, 3, 3,"], None),
    2: (["k_sweep1_group<8,", "k_sweep_group<8, 4, 4,"], None),
    3: (["k_sweep1_group<16,", "k_sweep_group<16, 5, 5,"], None),
    4: (["k_sweep1_group<32,", "k_sweep_group<32, 6, 6,"], None),
    5: None, 6: None, 7: None, 8: None,
    9: None, 10: None,
    11: (["k_sweep1_block<512,", "k_sweep_block<512, 13, 12,"], None),
    12: (["k_hub1_chunk", "k_hub1_final", "k_hub_count", "k_hub_scatter", "k_hub_agg",
          "k_hub_final"], ["k_hub1_final", "k_hub_final"]),
}
TIER_NAMES = {1: "k_sweep1_group<4> / k_sweep_group<4,3,3>", 2: "k_sweep1_group<8> / k_sweep_group<8,4,4>",
              3: "k_sweep1_group<16> / k_sweep_group<16,5,5>", 4: "k_sweep1_group<32> / k_sweep_group<32,6,6>",
              5: "k_sweep1_group<64> / k_sweep_group<64,7,7>", 6: "k_sweep1_group<64> / k_sweep_group<64,8,8>",
              7: "k_sweep1_group<64> / k_sweep_group<64,9,9>", 8: "k_sweep1_group<64> / k_sweep_group<64,10,10>",
              9: "k_sweep1_block<256> / k_sweep_block<256,11,11>",
              10: "k_sweep1_block<256> / k_sweep_block<256,12,11>",
              11: "k_sweep1_block<512> / k_sweep_block<512,13,12>",
              12: "first pass: k_hub1_chunk + k_hub1_final; later passes: k_hub_count + k_hub_scatter "
                  "+ k_hub_agg + k_hub_final"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "louvain_%s_hbm_traffic.txt")


def host_threads():
    """Threads this process can run at once: its CPU affinity, bounded by the
    cgroup CPU quota when one is set."""
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        pass
    quota = cgroup_cpus()
    if quota:
        n = min(n, max(1, int(quota + 0.999)))
    return max(1, n)


def cgroup_cpus():
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        return None if q == "max" else int(q) / int(per)
    except Exception:
        return None


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except Exception:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = None
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff,
            "cgroup_cpus": cgroup_cpus(), "threads_used": host_threads()}


def copy_probe(nbytes):
    """One hipMemcpy device-to-device of nbytes (a plain copy kernel)."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), ct.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ct.byref(b), ct.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ct.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ct.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def measured_traffic(workload_key, tier):
    """HBM bytes per launch of a sweep tier from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload: the tier's kernels' bytes
    (per-dispatch average x dispatches) over the tier's launches, or None."""
    path = TRAFFIC_PROFILE % workload_key
    spec = TIER_KERNELS.get(tier)
    if spec is None or not os.path.exists(path):
        return None
    per_kernel = {}
    with open(path) as f:
        for line in f:
            parts = line.split(None, 3)
            if line.startswith("#") or len(parts) < 4 or not parts[0].isdigit():
                continue
            per_kernel[parts[3].strip()] = (int(parts[0]), (float(parts[1]) + float(parts[2])) * 1e6)
    pats, markers = spec

    def lookup(pat):
        # every template instance of the kernel (entry layouts, table widths)
        return [v for k, v in per_kernel.items() if pat in k]

    total, launches = 0.0, 0
    for pat in pats:
        hits = lookup(pat)
        if not hits:
            return None
        total += sum(calls * nbytes for calls, nbytes in hits)
    for pat in (markers or pats):
        launches += sum(calls for calls, _ in lookup(pat))
    return total / launches if launches else None


def gen_batch_graphs(count, seed, threads):
    """The batching workload's SBM graphs (generated in parallel; the C++
    generator releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    from louvaingpu_amd import generators
    params = generators.batch_params(count, seed=seed)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda p: generators.sbm(p[0], 100, p[1], p[1] / 50, p[2]), params))


def run_batched(args):
    """4096 SBM graphs x 16 resolutions as graph sets (packed launches) against
    the plain per-graph launch loop of the same engine on a sample."""
    import oracle
    from louvaingpu_amd import Engine, GraphSet, generators
    count = args.batch_graphs
    gammas = generators.resolutions(16)
    nr = len(gammas)
    t0 = time.perf_counter()
    graphs = gen_batch_graphs(count, 3, host_threads())
    gen_s = time.perf_counter() - t0
    e_tot = sum(len(g[1]) for g in graphs)
    n_tot = sum(g[0] for g in graphs)
    print("[bench] batch: %d graphs, n=%d, e=%d, generated in %.1f s" % (count, n_tot, e_tot, gen_s),
          file=sys.stderr, flush=True)
    # chunks of consecutive graphs, bounded by union edges (graph edges x resolutions)
    budget = args.batch_chunk_edges
    chunks, cur, ce = [], [], 0
    for i, g in enumerate(graphs):
        if cur and ce + len(g[1]) * nr > budget:
            chunks.append(cur)
            cur, ce = [], 0
        cur.append(i)
        ce += len(g[1]) * nr
    if cur:
        chunks.append(cur)
    arena = int(budget * 200 + 400 * sum(graphs[i][0] for i in max(chunks, key=len)) * nr
                + (2 << 30))
    run_ms, load_s = 0.0, 0.0
    finals = {}
    sample = list(range(0, count, max(1, count // args.batch_sample)))
    sample_set = set(sample)
    with GraphSet(arena_bytes=arena) as gs:
        for ci, ch in enumerate(chunks):
            tl = time.perf_counter()
            gs.load([graphs[i] for i in ch], gammas)
            load_s += time.perf_counter() - tl
            ms = gs.run()
            run_ms += ms
            for j, gi in enumerate(ch):
                if gi in sample_set:
                    finals[gi] = [gs.final(j * nr + r) for r in range(nr)]
            print("[bench] batch chunk %d/%d: %d graphs, %.1f ms" % (ci + 1, len(chunks), len(ch), ms),
                  file=sys.stderr, flush=True)
        st = gs.engine.stats()
        arena_gb = (st.arena_capacity / 1e9, st.arena_high_water / 1e9)
    runs = count * nr
    # plain per-graph loop over the sample: one context per resolution, input resident
    engines = [Engine(gamma=g) for g in gammas]
    loop_ms, same = 0.0, True
    try:
        for gi in sample:
            g = graphs[gi]
            for r, eng in enumerate(engines):
                eng.load_edges(*g)
                tr = time.perf_counter()
                eng.run()
                loop_ms += (time.perf_counter() - tr) * 1e3
                same = same and bool(np.array_equal(eng.final_assignment(), finals[gi][r]))
    finally:
        for eng in engines:
            eng.close()
    sample_e = sum(len(graphs[i][1]) for i in sample)
    loop_est_s = loop_ms * 1e-3 * (e_tot / max(sample_e, 1))  # scaled by edges
    loop_est_runs_s = loop_ms * 1e-3 * (count / len(sample))
    # oracle on a few sampled graphs and resolutions
    oracle_ok = True
    for gi in sample[:4]:
        for r in (0, nr // 2, nr - 1):
            ro = oracle.run(*graphs[gi], gamma=gammas[r], threads=host_threads())
            oracle_ok = oracle_ok and bool(np.array_equal(ro.final_comm(), finals[gi][r]))
    batched_s = run_ms * 1e-3
    out = {
        "metric": "Louvain TEPS & end-to-end s at oracle modularity; batched vs per-graph",
        "value": runs / batched_s,
        "unit": "graph-runs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": run_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "sbm_batch_%dx%dgamma_seed3" % (count, nr), "graphs": count,
                   "resolutions": nr, "n_total": n_tot, "e_total": e_tot, "chunks": len(chunks),
                   "chunk_union_edges": budget, "parallelism": "single-gpu"},
        "batched_s": batched_s,
        "batched_load_s": load_s,
        "arena_gb": arena_gb[0], "arena_high_water_gb": arena_gb[1],
        "per_graph_loop_s_est": loop_est_runs_s,
        "per_graph_loop_s_est_by_edges": loop_est_s,
        "batch_vs_loop": loop_est_runs_s / batched_s,
        "loop_sample": "%d graphs x %d resolutions (every %d-th graph), %.1f s" % (
            len(sample), nr, max(1, count // args.batch_sample), loop_ms * 1e-3),
        "validated": same and oracle_ok,
        "identical_partitions_sample": same,
        "oracle_sample_ok": oracle_ok,
        "generation_s": gen_s,
    }
    return out


def run_serving(args):
    """Poisson stream (200 req/s) of SBM graphs drawn as in the batching
    workload plus 2% R-MAT scale-20 float graphs, served with continuous
    batching; p50/p99 latency, graphs/s, saturated throughput against the
    plain per-graph loop, and results checked against the request alone."""
    from concurrent.futures import ThreadPoolExecutor
    from louvaingpu_amd import Engine, Server, generators
    nreq = args.serve_requests
    rate = args.serve_rate
    rng = np.random.Generator(np.random.PCG64(4))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    is_rmat = rng.uniform(size=nreq) < 0.02
    params = generators.batch_params(nreq, seed=4)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=host_threads()) as ex:
        pool = list(ex.map(lambda sd: generators.rmat(20, 16, 0.57, 0.19, 0.19, sd, "float"),
                           range(4, 8)))
        sbms = list(ex.map(lambda p: generators.sbm(p[0], 100, p[1], p[1] / 50, p[2]), params))
    reqs = [pool[i % len(pool)] if is_rmat[i] else sbms[i] for i in range(nreq)]
    gen_s = time.perf_counter() - t0
    print("[bench] serve: %d requests (%d R-MAT 20), generated in %.1f s"
          % (nreq, int(is_rmat.sum()), gen_s), file=sys.stderr, flush=True)
    budget = args.serve_budget_gb << 30

    def stream(srv, times):
        ids = [None] * nreq
        start = time.perf_counter()
        for i in range(nreq):
            if times is not None:
                d = times[i] - (time.perf_counter() - start)
                if d > 0:
                    time.sleep(d)
            ids[i] = srv.submit(*reqs[i])
        return ids

    out = {}
    with Server(workers=args.serve_workers, budget_bytes=budget) as srv:
        srv.wait(srv.submit(*reqs[0]))  # warm the contexts
        ids = stream(srv, arrivals)
        res = [srv.wait(r) for r in ids]
        lat = np.array([r[1].t_done_ms - r[1].t_submit_ms for r in res])
        span = (max(r[1].t_done_ms for r in res) - min(r[1].t_submit_ms for r in res)) * 1e-3
        st = srv.stats()
        out.update({"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                    "mean_ms": float(lat.mean()), "graphs_per_s": nreq / span,
                    "sets": int(st.sets) - 1, "max_set": int(st.max_set)})
        # by class (the R-MAT 20 graphs are 30-60x the median SBM's edges), and
        # each latency split into the wait for a worker and the set's run
        wait = np.array([r[1].t_start_ms - r[1].t_submit_ms for r in res])
        for key, sel in (("sbm", ~is_rmat), ("rmat20", is_rmat)):
            if sel.any():
                out["by_class"] = out.get("by_class", {})
                out["by_class"][key] = {
                    "requests": int(sel.sum()),
                    "p50_ms": float(np.percentile(lat[sel], 50)),
                    "p99_ms": float(np.percentile(lat[sel], 99)),
                    "wait_p50_ms": float(np.percentile(wait[sel], 50)),
                    "wait_p99_ms": float(np.percentile(wait[sel], 99)),
                    "run_p50_ms": float(np.percentile((lat - wait)[sel], 50)),
                    "run_p99_ms": float(np.percentile((lat - wait)[sel], 99))}
        print("[bench] serve: p50 %.1f ms p99 %.1f ms" % (out["p50_ms"], out["p99_ms"]),
              file=sys.stderr, flush=True)
        # saturated: every request queued at once
        ts = time.perf_counter()
        ids = stream(srv, None)
        sat = [srv.wait(r) for r in ids]
        sat_s = time.perf_counter() - ts
        out["saturated_graphs_per_s"] = nreq / sat_s
        same_mix = all(np.array_equal(a[0], b[0]) for a, b in zip(res, sat))
    # plain per-graph loop (one context; each request uploaded, run, fetched) on a
    # sample, and alone == mix
    sample = list(range(0, nreq, max(1, nreq // args.serve_sample)))
    loop_s, alone_ok = 0.0, True
    with Engine() as eng:
        for i in sample:
            tr = time.perf_counter()  # a request end to end: upload, run, partition back
            eng.load_edges(*reqs[i])
            eng.run()
            fin = eng.final_assignment()
            loop_s += time.perf_counter() - tr
            alone_ok = alone_ok and bool(np.array_equal(fin, res[i][0]))
    loop_rate = len(sample) / loop_s
    out.update({
        "metric": "Louvain TEPS & end-to-end s at oracle modularity; batched vs per-graph",
        "value": out["graphs_per_s"], "unit": "graphs/s",
        "config": {"workload": "serve_poisson%d_sbm98_rmat20f2_seed4" % int(rate),
                   "requests": nreq, "rate_per_s": rate, "workers": args.serve_workers,
                   "budget_gb_per_worker": args.serve_budget_gb},
        "per_graph_loop_graphs_per_s": loop_rate,
        "saturated_vs_loop": out["saturated_graphs_per_s"] / loop_rate,
        "loop_sample": "%d requests" % len(sample),
        "alone_equals_mix": alone_ok and same_mix,
        "generation_s": gen_s,
    })
    return out


def workload(name):
    from louvaingpu_amd import generators
    if name == "rmat22":
        return "rmat22_ef16_int_seed1", generators.rmat22()
    if name == "sbm10k":
        return "sbm_10k_100blocks_seed0", generators.sbm_10k()
    if name == "rmat26f":
        return "rmat26_ef16_float_seed2", generators.rmat26_float()
    if name.startswith("rmat"):
        sc = int(name[4:])
        return "rmat%d_ef16_int_seed1" % sc, generators.rmat(sc, 16, 0.57, 0.19, 0.19, 1, "int")
    raise SystemExit("unknown workload " + name)


def e2e_runs(eng, n, hsrc, hdst, hw, steps, warmup, label):
    """End-to-end steps (pinned edge list -> device -> run -> every level's
    Partition and the final partition in pinned host buffers; the Partitions
    are copied while later levels run, lv_run_partitions).  Returns (seconds
    per step, per-part ms, levels); 'run' includes the overlapped copies,
    'results' what is left after the run."""
    from louvaingpu_amd.engine import pinned_empty
    outs, fin = None, None
    parts = np.zeros(3)
    levels = 0
    total = 0.0
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        eng.load_edges(n, hsrc, hdst, hw)
        t1 = time.perf_counter()
        if outs is None:  # the first run finds the level count
            levels = eng.run()
            t2 = time.perf_counter()
            for l in range(levels):
                eng.level_partition(l)
            eng.final_assignment()
        else:
            levels = eng.run_partitions([o[0] for o in outs], [o[1] for o in outs], fin)
            t2 = time.perf_counter()
        t3 = time.perf_counter()
        if i >= warmup:
            parts += (t1 - t0, t2 - t1, t3 - t2)
            total += t3 - t0
        else:
            print("[bench] %s e2e warmup %d: %.2f ms (load %.2f, run %.2f, results %.2f)"
                  % (label, i, (t3 - t0) * 1e3, (t1 - t0) * 1e3, (t2 - t1) * 1e3,
                     (t3 - t2) * 1e3), file=sys.stderr, flush=True)
        if outs is None:  # page-locked result buffers for the level count of the first run
            outs = [(pinned_empty(n, np.int32), pinned_empty(n, np.int32)) for l in range(levels)]
            fin = pinned_empty(n, np.int32)
    k = max(steps, 1)
    return total / k, {"load_ms": parts[0] * 1e3 / k, "run_ms": parts[1] * 1e3 / k,
                       "results_ms": parts[2] * 1e3 / k}, levels


def measure_graph(wname, graph, steps, warmup, traffic_key, fanout=0):
    """End-to-end and device-resident timing of one graph, with the dominant
    sweep tier's roofline.  Returns the measurement dict and the engine's final
    partition and modularity."""
    from louvaingpu_amd import Engine
    from louvaingpu_amd._lib import LV_SWEEP_TIERS
    from louvaingpu_amd.engine import pinned_empty
    n, src, dst, w = graph
    e = len(src)
    nnz0 = 2 * e
    # the caller's edge list in page-locked host memory (filled outside the steps)
    hsrc, hdst, hw = pinned_empty(e, np.int32), pinned_empty(e, np.int32), pinned_empty(e, w.dtype)
    hsrc[:], hdst[:], hw[:] = src, dst, w
    with Engine(fanout=fanout) as eng:
        e2e_s, e2e_parts, levels = e2e_runs(eng, n, hsrc, hdst, hw, steps, warmup, wname)
        # the same runs with the input already resident in HBM (the last load's)
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.run()
        dev_ms = (time.perf_counter() - t0) * 1000.0 / max(steps, 1)
    del hsrc, hdst, hw
    # phases and the roofline: HIP events around every phase and tier launch
    eng = Engine(timing=True)
    try:
        eng.load_edges(n, src, dst, w)
        for i in range(warmup):
            eng.run()
        eng.reset_stats()
        for _ in range(steps):
            eng.run()
        st = eng.stats()
        infos = [eng.level_info(l) for l in range(eng.num_levels())]
        final = eng.final_assignment()
        q_gpu = eng.final_modularity()
        arena = (st.arena_capacity / 1e9, st.arena_high_water / 1e9)
    finally:
        eng.close()
    tier_ms = [st.tier_ms[t] for t in range(LV_SWEEP_TIERS)]
    dom = int(np.argmax(tier_ms))
    launches = max(1, st.tier_launches[dom])
    avg_ms = st.tier_ms[dom] / launches
    bytes_per_launch = st.tier_bytes[dom] / launches
    cand_per_launch = st.tier_cand[dom] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    achieved_p0 = ((bytes_per_launch - 8.0 * cand_per_launch) / (avg_ms * 1e-3) / 1e9
                   if avg_ms > 0 else 0.0)
    k = max(steps, 1)
    phases = {"csr": st.csr_ms / k, "init": st.init_ms / k, "sweep": st.sweep_ms / k,
              "apply": st.apply_ms / k, "renumber": st.renumber_ms / k,
              "coarsen": st.coarsen_ms / k, "flatten": st.flatten_ms / k}
    m = {
        "config": {"workload": wname, "n": n, "e": e, "nnz": nnz0, "levels": levels,
                   "passes": [i.passes for i in infos], "level_n": [i.n for i in infos],
                   "level_nnz": [i.nnz for i in infos], "parallelism": "single-gpu",
                   "fanout": fanout},
        "value": nnz0 / e2e_s,
        "ms_per_step": e2e_s * 1e3,
        "e2e_parts_ms": e2e_parts,
        "device_resident_ms": dev_ms,
        "device_resident_teps": nnz0 / (dev_ms * 1e-3),
        "roofline": {
            "bound": "hbm",
            "kernel": "sweep tier %d (%s)" % (dom + 1, TIER_NAMES[dom + 1]),
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_key, dom + 1),
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
            "candidates_per_launch": cand_per_launch,
            "frac_p0": achieved_p0 / HBM_PEAK_GBS,
        },
        "phase_ms_per_step": phases,
        "coarsen_fraction": phases["coarsen"] / dev_ms if dev_ms > 0 else None,
        "sweep_gbs": st.sweep_bytes / (st.sweep_ms * 1e-3) / 1e9 if st.sweep_ms > 0 else 0.0,
        "sweep_frac": (st.sweep_bytes / (st.sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                       if st.sweep_ms > 0 else 0.0),
        "sweep_teps": st.sweep_entries / (st.sweep_ms * 1e-3) if st.sweep_ms > 0 else 0.0,
        "tier_ms_per_step": [t / k for t in tier_ms],
        "arena_gb": arena[0], "arena_high_water_gb": arena[1],
        "modularity": q_gpu, "communities": int(final.max()) + 1,
    }
    return m, final, q_gpu


def cpu_baselines(n, src, dst, w, final, q_gpu, wname, nnz0, gpu_s):
    """The oracle on every usable thread (full graph) and on one thread (R-MAT
    scale-20 sample)."""
    import oracle
    from louvaingpu_amd import generators
    threads = host_threads()
    c0 = time.perf_counter()
    r = oracle.run(n, src, dst, w, threads=threads)
    cpu_s = time.perf_counter() - c0
    same = bool(np.array_equal(r.final_comm(), final))
    q_or = r.final_modularity()
    r.close()
    sn, ss, sd, sw = generators.rmat(20, 16, 0.57, 0.19, 0.19, 1, "int")
    c0 = time.perf_counter()
    r1 = oracle.run(sn, ss, sd, sw, threads=1)
    one_s = time.perf_counter() - c0
    r1.close()
    return {
        "cpu_baseline": {
            "value": nnz0 / cpu_s, "unit": "TEPS", "cores": threads, "kind": "port",
            "sample": "oracle (C restatement, OpenMP, %d threads) full multi-level run on the "
                      "same %s graph, %.1f s" % (threads, wname, cpu_s)},
        "cpu_baseline_1thread": {
            "value": 2 * len(ss) / one_s, "unit": "TEPS", "cores": 1, "kind": "port",
            "sample": "oracle on one thread, full multi-level run on R-MAT scale 20 (edge factor "
                      "16, .57/.19/.19/.05, int weights, seed 1; %d entries), %.1f s"
                      % (2 * len(ss), one_s)},
        "host_cpu": host_cpu(),
        "validated": same and abs(q_or - q_gpu) <= 1e-9 * max(1.0, abs(q_or)),
        "oracle_modularity": q_or,
        "speedup_vs_cpu": cpu_s / gpu_s,
    }


def rmat26f_object(args):
    """The R-MAT scale-26 float graph: GPU side only, validated against the
    committed oracle digest of its final partition."""
    import hashlib
    from louvaingpu_amd import generators
    t0 = time.perf_counter()
    graph = generators.rmat26_float()
    gen_s = time.perf_counter() - t0
    print("[bench] rmat26f generated in %.1f s" % gen_s, file=sys.stderr, flush=True)
    m, final, q = measure_graph("rmat26_ef16_float_seed2", graph, args.rmat26_steps, 1, "rmat26f",
                                args.fanout)
    with open(os.path.join(ROOT, "tests", "golden", "rmat26f_oracle.json")) as f:
        fx = json.load(f)
    m["validated"] = (hashlib.sha256(np.ascontiguousarray(final).tobytes()).hexdigest()
                      == fx["final_sha256"] and
                      abs(q - float.fromhex(fx["final_modularity"])) <= 1e-9)
    m["oracle_s_16_threads"] = fx["oracle_s"]
    m["speedup_vs_cpu"] = fx["oracle_s"] / (m["ms_per_step"] * 1e-3)
    m["steps"] = args.rmat26_steps
    return m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / validation")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against a known byte count)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-per-graph measurement of the default run")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving measurement of the default run")
    ap.add_argument("--no-rmat26", action="store_true",
                    help="skip the R-MAT scale-26 object of the default run")
    ap.add_argument("--rmat26-steps", type=int, default=3)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-workers", type=int, default=2)
    ap.add_argument("--serve-budget-gb", type=int, default=24)
    ap.add_argument("--serve-sample", type=int, default=200)
    ap.add_argument("--batch-sample", type=int, default=64,
                    help="graphs of the per-graph loop sample (each at all resolutions)")
    ap.add_argument("--batch-chunk-edges", type=int, default=300_000_000,
                    help="union edges (graph edges x resolutions) per graph set")
    ap.add_argument("--fanout", type=int, default=0, choices=(0, 1, 2),
                    help="lv_config.fanout of the end-to-end and device-resident runs "
                         "(1 = one stream, so a kernel trace shows each kernel's own time)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2

    if args.workload == "batch":
        print(json.dumps(run_batched(args)), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(run_serving(args)), flush=True)
        return 0

    wname, graph = workload(args.workload)
    n, src, dst, w = graph
    nnz0 = 2 * len(src)
    print("[bench] workload %s n=%d e=%d nnz=%d" % (wname, n, len(src), nnz0), file=sys.stderr,
          flush=True)
    m, final, q_gpu = measure_graph(wname, graph, args.steps, args.warmup, args.workload,
                                    args.fanout)
    out = {
        "metric": "Louvain TEPS & end-to-end s at oracle modularity; batched vs per-graph",
        "value": m.pop("value"),
        "unit": "TEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": m.pop("config"),
        "roofline": m.pop("roofline"),
    }
    out["e2e_s"] = out["ms_per_step"] * 1e-3
    out.update(m)
    if not args.no_cpu:
        out.update(cpu_baselines(n, src, dst, w, final, q_gpu, wname, nnz0, out["e2e_s"]))
    else:
        out["cpu_baseline"] = None
        out["validated"] = None
    del graph, src, dst, w
    if args.copy_probe_mb:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mb << 20)
    if args.workload == "rmat22" and not args.no_rmat26:
        out["rmat26f"] = rmat26f_object(args)
    if not args.no_batch:
        b = run_batched(args)
        out["batched"] = {k: b[k] for k in (
            "value", "unit", "config", "batched_s", "batched_load_s", "arena_high_water_gb",
            "per_graph_loop_s_est", "batch_vs_loop", "loop_sample", "validated")}
    if not args.no_serve:
        v = run_serving(args)
        out["serving"] = {k: v[k] for k in (
            "config", "p50_ms", "p99_ms", "by_class", "graphs_per_s", "saturated_graphs_per_s",
            "per_graph_loop_graphs_per_s", "saturated_vs_loop", "alone_equals_mix", "max_set")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
