#!/usr/bin/env python3
"""Louvain benchmark: TEPS of the full multi-level Louvain on one MI355X.

One step = one complete run of the reference's pipeline (CSR build, init,
local-moving sweeps, renumbering, coarsening at every level, flattening) on
the R-MAT scale-22 graph (edge factor 16, a/b/c/d = .57/.19/.19/.05,
duplicates merged into int32 weights, seed 1), with the edge list already
resident in HBM.  value = TEPS = directed CSR entries of the input (2E) /
seconds per step.

Besides the JSON line's required fields it reports:
  roofline     the dominant local-moving sweep tier (one kernel, or the four
               kernels of the hub pipeline): algorithmic bytes 12*nnz + 24*N per
               launch (community-total reads P not counted) / its mean HIP-event
               duration, against 8 TB/s; traffic = measured HBM bytes per launch
               from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
  cpu_baseline the oracle (C restatement of the reference semantics, OpenMP)
               on the same graph, timed on this host; its partition is also
               the validation of the GPU result
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
# kernel, so their traffic is not separable.
TIER_KERNELS = {
    1: (["k_sweep1_group<4,", "k_sweep_group<4, 3, 3,"], None),
    2: (["k_sweep1_group<8,", "k_sweep_group<8, 4, 4,"], None),
    3: (["k_sweep1_group<16,", "k_sweep_group<16, 5, 5,"], None),
    4: (["k_sweep1_group<32,", "k_sweep_group<32, 6, 6,"], None),
    5: None, 6: None, 7: None, 8: None,
    9: (["k_sweep1_block<256,", "k_sweep_block<256, 12, 11,"], None),
    10: (["k_hub1_chunk", "k_hub1_final", "k_hub_count", "k_hub_scatter", "k_hub_agg",
          "k_hub_final"], ["k_hub1_final", "k_hub_final"]),
}
TIER_NAMES = {1: "k_sweep1_group<4> / k_sweep_group<4,3,3>", 2: "k_sweep1_group<8> / k_sweep_group<8,4,4>",
              3: "k_sweep1_group<16> / k_sweep_group<16,5,5>", 4: "k_sweep1_group<32> / k_sweep_group<32,6,6>",
              5: "k_sweep1_group<64> / k_sweep_group<64,7,7>", 6: "k_sweep1_group<64> / k_sweep_group<64,8,8>",
              7: "k_sweep1_group<64> / k_sweep_group<64,9,9>", 8: "k_sweep1_group<64> / k_sweep_group<64,10,10>",
              9: "k_sweep1_block<256> / k_sweep_block<256,12,11>",
              10: "first pass: k_hub1_chunk + k_hub1_final; later passes: k_hub_count + k_hub_scatter "
                  "+ k_hub_agg + k_hub_final"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "louvain_%s_hbm_traffic.txt")


def host_threads():
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return max(1, min(n, 16))


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
        hits = [v for k, v in per_kernel.items() if pat in k]
        return hits[0] if len(hits) == 1 else None

    total, launches = 0.0, 0
    for pat in pats:
        hit = lookup(pat)
        if hit is None:
            return None
        total += hit[0] * hit[1]
    for pat in (markers or pats):
        launches += lookup(pat)[0]
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
        lat_r = lat[is_rmat]
        if len(lat_r):
            out["p50_ms_rmat20"] = float(np.percentile(lat_r, 50))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / validation")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against a known byte count)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-per-graph measurement of the default run")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving measurement of the default run")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-workers", type=int, default=2)
    ap.add_argument("--serve-budget-gb", type=int, default=24)
    ap.add_argument("--serve-sample", type=int, default=200)
    ap.add_argument("--batch-sample", type=int, default=64,
                    help="graphs of the per-graph loop sample (each at all resolutions)")
    ap.add_argument("--batch-chunk-edges", type=int, default=300_000_000,
                    help="union edges (graph edges x resolutions) per graph set")
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
    from louvaingpu_amd import Engine
    from louvaingpu_amd._lib import LV_SWEEP_TIERS

    wname, (n, src, dst, w) = workload(args.workload)
    e = len(src)
    nnz0 = 2 * e
    print("[bench] workload %s n=%d e=%d nnz=%d" % (wname, n, e, nnz0), file=sys.stderr, flush=True)

    eng = Engine(timing=True)
    t_load0 = time.perf_counter()
    eng.load_edges(n, src, dst, w)
    t_load = time.perf_counter() - t_load0
    for i in range(args.warmup):
        eng.run()
        print("[bench] warmup %d: %.2f ms" % (i, eng.last_run_ms()), file=sys.stderr, flush=True)
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / max(args.steps, 1)
    st = eng.stats()
    levels = eng.num_levels()
    infos = [eng.level_info(l) for l in range(levels)]
    final = eng.final_assignment()
    q_gpu = eng.final_modularity()

    # PCIe-inclusive: host edge list -> device -> dendrogram and final partition on the host
    tp0 = time.perf_counter()
    eng.load_edges(n, src, dst, w)
    eng.run()
    for l in range(levels):
        eng.level_partition(l)
    eng.final_assignment()
    pcie_s = time.perf_counter() - tp0

    # dominant sweep kernel (tier with the largest device time)
    tier_ms = [st.tier_ms[t] for t in range(LV_SWEEP_TIERS)]
    dom = int(np.argmax(tier_ms))
    launches = max(1, st.tier_launches[dom])
    avg_ms = st.tier_ms[dom] / launches
    bytes_per_launch = st.tier_bytes[dom] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    sweep_gbs = st.sweep_bytes / (st.sweep_ms * 1e-3) / 1e9 if st.sweep_ms > 0 else 0.0
    sweep_teps = st.sweep_entries / (st.sweep_ms * 1e-3) if st.sweep_ms > 0 else 0.0
    steps = max(args.steps, 1)

    out = {
        "metric": "Louvain TEPS & end-to-end s at oracle modularity; batched vs per-graph",
        "value": nnz0 / (ms * 1e-3),
        "unit": "TEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": wname,
            "n": n, "e": e, "nnz": nnz0,
            "levels": levels,
            "passes": [i.passes for i in infos],
            "level_n": [i.n for i in infos],
            "level_nnz": [i.nnz for i in infos],
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "sweep tier %d (%s)" % (dom + 1, TIER_NAMES[dom + 1]),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.workload, dom + 1),
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
        },
        "e2e_s": ms * 1e-3,
        "pcie_inclusive_s": pcie_s,
        "load_s": t_load,
        "phase_ms_per_step": {
            "csr": st.csr_ms / steps, "init": st.init_ms / steps, "sweep": st.sweep_ms / steps,
            "apply": st.apply_ms / steps, "renumber": st.renumber_ms / steps,
            "coarsen": st.coarsen_ms / steps, "flatten": st.flatten_ms / steps,
        },
        "sweep_gbs": sweep_gbs,
        "sweep_teps": sweep_teps,
        "tier_ms_per_step": [t / steps for t in tier_ms],
        "arena_gb": st.arena_capacity / 1e9,
        "arena_high_water_gb": st.arena_high_water / 1e9,
        "modularity": q_gpu,
        "communities": int(final.max()) + 1,
    }

    if not args.no_cpu:
        import oracle
        threads = args.cpu_threads or host_threads()
        c0 = time.perf_counter()
        r = oracle.run(n, src, dst, w, threads=threads)
        cpu_s = time.perf_counter() - c0
        same = bool(np.array_equal(r.final_comm(), final))
        q_or = r.final_modularity()
        out["cpu_baseline"] = {
            "value": nnz0 / cpu_s,
            "unit": "TEPS",
            "cores": threads,
            "kind": "port",
            "sample": "oracle (C restatement, OpenMP) full multi-level run on the same %s graph, "
                      "%.1f s" % (wname, cpu_s),
        }
        out["validated"] = same and abs(q_or - q_gpu) <= 1e-9 * max(1.0, abs(q_or))
        out["oracle_modularity"] = q_or
        out["speedup_vs_cpu"] = cpu_s / (ms * 1e-3)
    else:
        out["cpu_baseline"] = None
        out["validated"] = None
    if args.copy_probe_mb:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mb << 20)
    eng.close()
    if not args.no_batch:
        b = run_batched(args)
        out["batched"] = {k: b[k] for k in (
            "value", "unit", "config", "batched_s", "batched_load_s", "per_graph_loop_s_est",
            "batch_vs_loop", "loop_sample", "validated")}
    if not args.no_serve:
        v = run_serving(args)
        out["serving"] = {k: v[k] for k in (
            "config", "p50_ms", "p99_ms", "graphs_per_s", "saturated_graphs_per_s",
            "per_graph_loop_graphs_per_s", "saturated_vs_loop", "alone_equals_mix", "max_set")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
