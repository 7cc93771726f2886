"""Benchmark: PageRank GTEPS per iteration on one MI355X.

Workload (default): R-MAT scale 24 (16.8M nodes, ~268M directed edges drawn,
a/b/c/d = .57/.19/.19/.05, seed 11), self-loops and duplicates removed,
edge weights log-normal(0, 1) stored as bf16; pagerank_weighted_run with
damping 0.85, tolerance 1e-12, at most 200 iterations; f32 delta sweeps with
f64 anchor sweeps, f64 residual / dangling sums.  One step = one complete
pagerank_weighted_run with the graph already resident in HBM (results left
on the device).  value = edges * sweeps / seconds (GTEPS per iteration).

Also reported: the roofline of the dominant kernel (the f32 delta sweep:
algorithmic bytes (4 + 2) * m + 16 * n + 4 per sweep against 8 TB/s,
durations from HIP events on the library's stream around its launches), the error against
the f64 oracle, the oracle's single-core CPU baseline on a bounded sample of
sweeps, and two more legs on an R-MAT scale-22 graph: batched personalized
PageRank (1024 queries) against the plain one-query-per-run path, and a
serving run with Poisson arrivals.

Further legs on the same resident graph: its device and host builds, and
Katz / eigenvector / HITS centrality (f64 power iterations).

--copy-probe BYTES runs a plain device copy of BYTES instead (the
calibration of the FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")

CONFIGS = {
    # scale, seed, weights, storage, tol, max_it, w_bytes
    "c3": dict(scale=24, seed=11, weights="lognormal", p0=0.0, p1=1.0, rounding="bf16",
               storage="bf16", tol=1e-12, max_it=200, w_bytes=2,
               workload="rmat24-bf16-lognormal pagerank_weighted d0.85 tol1e-12"),
    "c2": dict(scale=20, seed=7, weights="uniform", p0=0.1, p1=1.0, rounding="f32",
               storage="f32", tol=1e-9, max_it=200, w_bytes=4,
               workload="rmat20-f32-uniform pagerank_weighted d0.85 tol1e-9"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(e, cfg, n_sweeps, damping=0.85):
    """The oracle's faithful push sweep (pagerank.rs:239-274), one core, on a
    bounded sample of sweeps of the same workload."""
    import oracle

    info = machine_info()

    ws = oracle.out_wsum(e.offsets, e.weights)
    x = np.full(e.n, 1.0 / e.n)
    t0 = time.perf_counter()
    for _ in range(n_sweeps):
        x, _ = oracle.push_sweep(e.offsets, e.targets, x, damping, e.weights, ws)
    dt = time.perf_counter() - t0
    return {"value": round(e.m * n_sweeps / dt / 1e9, 4), "unit": "GTEPS", "cores": 1,
            "kind": "port", "cpu_model": info["cpu_model"], "nproc": info["nproc"],
            "cores_usable": info["cores_usable"],
            "sample": f"{n_sweeps} push sweeps of the f64 oracle (single thread) on the same "
                      f"graph, {dt:.1f} s"}


def oracle_reference(e, tol, max_it):
    """The f64 result of pagerank_weighted_run by the oracle's pull
    restatement (bit-identical to the push form; OpenMP over rows), and the
    same iteration carried in extended precision (the exact iteration to
    ~1e-19), from which the f64 reference drifts by its own rounding."""
    import oracle

    in_off, in_src, in_w = oracle.transpose(e.offsets, e.targets, e.weights)
    ws = oracle.out_wsum(e.offsets, e.weights)
    t0 = time.perf_counter()
    f64 = oracle.pull_run(1, in_off, in_src, ws, w=in_w, damping=0.85, max_iterations=max_it,
                          tolerance=tol, threads=0)
    f64_s = time.perf_counter() - t0
    ld = oracle.pull_run_ld(1, in_off, in_src, w=in_w, damping=0.85, max_iterations=max_it,
                            tolerance=tol, threads=0)
    return f64, ld, f64_s


def top_agreement(a, b, k=100):
    """|top-k(a) & top-k(b)| / k (ids by descending score, ties by id)."""
    ta = set(np.argsort(-a, kind="stable")[:k].tolist())
    tb = set(np.argsort(-b, kind="stable")[:k].tolist())
    return len(ta & tb) / float(k)


def machine_info():
    """The host the CPU figures ran on: model, cores visible, cores this
    process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", usable) or usable)
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_usable": usable,
            "omp_threads": threads}


def copy_peak(nbytes=1 << 30, reps=5):
    """Measured device copy bandwidth: hipMemcpyAsync of nbytes device to
    device (the runtime's copy kernel), HIP events around it; read + write
    bytes over the best of reps."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    a, b = vp(), vp()
    e0, e1 = vp(), vp()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        best = None
        for _ in range(reps):
            hip.hipEventRecord(e0, None)
            assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0  # D2D
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
            best = ms.value if best is None else min(best, ms.value)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    gbs = 2 * nbytes / (best / 1e3) / 1e9
    return {"what": "device-to-device copy of 1 GiB (hipMemcpyAsync), read + write bytes over "
                    "the best of 5, HIP events",
            "gbs": round(gbs, 1), "frac_of_nominal": round(gbs / HBM_PEAK_GBS, 3)}


def c1_leg(ga, generators, reps=20):
    """C1: directed G(2000, 0.005) from seed 42 as a dense AdjacencyMatrix,
    pagerank_run at d 0.85, tol 1e-10, max 100 (pagerank.rs:122-187): the GPU
    (graph resident, scores downloaded each run) against the oracle, whose
    CPU time includes the dense adapter's O(n^2) neighbour materialisation
    (graph.rs:128-135)."""
    import oracle

    e = generators.erdos_renyi(2000, 0.005, 42)
    n = e.n
    mat = np.zeros((n, n))
    for u in range(n):
        mat[u, e.targets[e.offsets[u]:e.offsets[u + 1]]] = 1.0
    from graphops_amd import _lib as L

    cfg = ga.PageRankConfig(0.85, 100, 1e-10)
    ctx = ga.Context(0)
    dg = ga.DeviceGraph.from_dense(mat, weighted=False, ctx=ctx)
    launch_ms = {}
    for name, mode in (("queued", L.LAUNCH_QUEUED), ("graph", L.LAUNCH_GRAPH)):
        ctx.set_option(L.OPT_LAUNCH, mode)
        r = ga.pagerank_run(dg, cfg, ctx=ctx)
        t0 = time.perf_counter()
        for _ in range(reps):
            r = ga.pagerank_run(dg, cfg, ctx=ctx)
        launch_ms[name] = (time.perf_counter() - t0) / reps * 1e3
    ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_AUTO)
    gpu_s = min(launch_ms.values()) / 1e3
    t0 = time.perf_counter()
    off, nb, _ = oracle.dense_neighbors(mat)
    ref = oracle.pagerank_run(off, nb, 0.85, 100, 1e-10)
    cpu_s = time.perf_counter() - t0
    d = np.abs(r.scores - ref.scores)
    m = int(off[-1])
    dg.close()
    ctx.close()
    return {"workload": "er2000 p0.005 seed42 dense AdjacencyMatrix, pagerank_run d0.85 tol1e-10",
            "n": n, "m": m, "sweeps": r.iterations, "oracle_sweeps": ref.iterations,
            "gpu_ms_per_run": round(gpu_s * 1e3, 3),
            "launch_ms_per_run": {k: round(v, 3) for k, v in launch_ms.items()},
            "gpu_gteps": round(m * r.iterations / gpu_s / 1e9, 3),
            "accuracy": {"l1_vs_f64_oracle": float(d.sum()), "max_abs": float(d.max()),
                         "top100_agreement": top_agreement(r.scores, ref.scores)},
            "cpu_baseline": {"value": round(cpu_s * 1e3, 2), "unit": "ms per run", "cores": 1,
                             "kind": "port",
                             "sample": "the oracle's pagerank_run with the dense adapter's "
                                       "neighbour lists built from the matrix (one run)"},
            "speedup_vs_cpu": round(cpu_s / gpu_s, 1)}


def c2_leg(ga, generators, reps=10, push_sweeps=3, find_edge_frac=0.02):
    """C2: R-MAT scale 20, weights U[0.1, 1) stored f32, seed 7,
    pagerank_weighted_run at d 0.85, tol 1e-9, max 200 (pagerank.rs:210-289):
    GTEPS per iteration, the f32 sweep's roofline, accuracy against the f64
    oracle and the extended-precision iteration; CPU: the oracle's push sweep
    on one core, its OpenMP pull on every usable core, and the reference's
    per-edge find_edge lookup (graph.rs:205-211) timed on a sample of
    sources."""
    import oracle

    c = CONFIGS["c2"]
    e = generators.rmat(c["scale"], 16, 0.57, 0.19, 0.19, c["seed"], c["weights"], c["p0"],
                        c["p1"], c["rounding"])
    n, m = e.n, e.m
    cfg = ga.PageRankConfig(0.85, c["max_it"], c["tol"])
    ctx = ga.Context(0, timing=True)
    dg = ga.DeviceGraph.from_edges(e, weight_storage=c["storage"], ctx=ctx)
    ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
    sweeps = 0
    f32_ms = 0.0
    n32 = 0
    t0 = time.perf_counter()
    for _ in range(reps):
        r = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
        sweeps += r.iterations
        f32_ms += r.stats["fp32_sweep_ms"]
        n32 += r.stats["sweeps_fp32"]
    wall = time.perf_counter() - t0
    scores = ga.pagerank_weighted_run(dg, cfg, ctx=ctx).scores
    delta_sweep = "blocked" if r.stats["sweeps_blocked"] else "pull"
    # the pull's row plan and its autotuner (gops_graph_tune_rows): the same
    # runs as pulls, with the default plan and the tuned one
    from graphops_amd import _lib as L

    def pull_runs():
        n_sw, f32, n32_ = 0, 0.0, 0
        t = time.perf_counter()
        for _ in range(reps):
            rp = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
            n_sw += rp.iterations
            f32 += rp.stats["fp32_sweep_ms"]
            n32_ += rp.stats["sweeps_fp32"]
        return n_sw, f32, n32_, time.perf_counter() - t

    ctx.set_option(L.OPT_DELTA_SWEEP, L.DELTA_SWEEP_PULL)
    try:
        p_sweeps, p_f32, p_n32, p_wall = pull_runs()
        tune = ga.graph_tune_rows(dg)
        t_sweeps, t_f32, t_n32, t_wall = pull_runs()
    finally:
        ctx.set_option(L.OPT_DELTA_SWEEP, L.DELTA_SWEEP_AUTO)
    tune.update({"delta_sweep": "pull",
                 "default_plan_gteps": round(m * p_sweeps / p_wall / 1e9, 2),
                 "default_plan_f32_sweep_us_in_runs": round(p_f32 / max(p_n32, 1) * 1e3, 2),
                 "gteps": round(m * t_sweeps / t_wall / 1e9, 2),
                 "f32_sweep_us_in_runs": round(t_f32 / max(t_n32, 1) * 1e3, 2)})
    log(f"c2 row plan: {tune}")
    dg.close()
    ctx.close()
    alg = (4 + c["w_bytes"]) * m + 16 * n + 4
    avg_f32 = f32_ms / max(n32, 1) / 1e3
    in_off, in_src, in_w = oracle.transpose(e.offsets, e.targets, e.weights)
    ws = oracle.out_wsum(e.offsets, e.weights)
    t0 = time.perf_counter()
    ref = oracle.pull_run(1, in_off, in_src, ws, w=in_w, damping=0.85, max_iterations=c["max_it"],
                          tolerance=c["tol"], threads=0)
    all_core_s = time.perf_counter() - t0
    exact = oracle.pull_run_ld(1, in_off, in_src, w=in_w, damping=0.85,
                               max_iterations=c["max_it"], tolerance=c["tol"], threads=0)
    d = np.abs(scores - ref.scores)
    x = np.full(n, 1.0 / n)
    t0 = time.perf_counter()
    for _ in range(push_sweeps):
        x, _ = oracle.push_sweep(e.offsets, e.targets, x, 0.85, e.weights, ws)
    push_s = (time.perf_counter() - t0) / push_sweeps
    # find_edge: the scatter of a random sample of sources with each entry's
    # weight found by a scan of its source's list; extrapolated to every
    # source by the sum of squared out-degrees (the scan cost)
    deg = np.diff(e.offsets.astype(np.int64))
    rng = np.random.default_rng(3)
    us = np.sort(rng.choice(n, size=max(1, int(find_edge_frac * n)), replace=False)).astype(np.uint32)
    t0 = time.perf_counter()
    oracle.push_find_edge(us, e.offsets, e.targets, e.weights, ws, np.full(n, 1.0 / n))
    fe_s = time.perf_counter() - t0
    scale_up = float((deg.astype(np.float64) ** 2).sum() / max((deg[us].astype(np.float64) ** 2).sum(), 1.0))
    info = machine_info()
    return {"workload": c["workload"], "n": n, "m": m, "sweeps_per_run": r.iterations,
            "gteps": round(m * sweeps / wall / 1e9, 2), "ms_per_run": round(wall / reps * 1e3, 3),
            "delta_sweep": delta_sweep,
            "f32_sweep_us": round(avg_f32 * 1e6, 2),
            "row_plan_autotune": tune,
            "roofline": {"bound": "hbm", "alg_bytes_per_launch": alg,
                         "achieved": round(alg / avg_f32 / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / avg_f32 / 1e9 / HBM_PEAK_GBS, 4)},
            "accuracy": {"l1_vs_f64_oracle": float(d.sum()), "max_abs": float(d.max()),
                         "top100_agreement": top_agreement(scores, ref.scores),
                         "oracle_iterations": ref.iterations, "gpu_iterations": r.iterations,
                         "l1_gpu_vs_exact": float(np.abs(scores - exact.scores).sum()),
                         "exact_iterations": exact.iterations},
            "cpu_baseline": {"value": round(m / push_s / 1e9, 4), "unit": "GTEPS", "cores": 1,
                             "kind": "port",
                             "sample": f"{push_sweeps} push sweeps of the f64 oracle, one thread "
                                       "(adjacency lists with the weights beside the targets)"},
            "best_cpu": {"value": round(m * ref.iterations / all_core_s / 1e9, 4),
                         "unit": "GTEPS", "cores": info["omp_threads"], "kind": "port",
                         "sample": f"the oracle's OpenMP pull, the whole run ({ref.iterations} "
                                   "sweeps)"},
            "find_edge_footnote": {
                "what": "the reference's per-entry edge_weight through petgraph's find_edge "
                        "(graph.rs:205-211, pagerank.rs:257): a scan of the source's list",
                "sample_sources": int(us.size), "sample_s": round(fe_s, 3),
                "projected_s_per_sweep": round(fe_s * scale_up, 2),
                "ns_per_edge": round(fe_s * scale_up / m * 1e9, 1),
                "cores": 1}}


PPR_GRAPH = dict(scale=22, seed=13)  # unit weights


def ppr_batch_leg(ga, generators, e, dg, nq=1024, plain_sample=32, cpu_queries=3):
    """Batched personalized PageRank: nq queries, |S| ~ U{1..64} seeds drawn
    proportional to out-degree (the source of a uniformly drawn edge),
    p-hat uniform on S, d 0.85, tol 1e-8; the tile width from the library's
    autotuner.  Compared with the plain one-query-per-run path on a sample
    of the same queries, and with the oracle (f64) on a few of them."""
    import oracle

    cfg = ga.PageRankConfig(0.85, 200, 1e-8)
    off, seeds = generators.seed_sets(e, nq, 1, 64, log_uniform=False, seed=5)
    tile, costs = ga.ppr_tile_info(dg, nq)
    log(f"ppr: autotuned tile {tile}, ms per query-sweep {costs}")
    ga.ppr_batch_run(dg, cfg, off[:65], seeds[:off[64]], with_scores=False)  # warm-up
    t0 = time.perf_counter()
    _, runs = ga.ppr_batch_run(dg, cfg, off, seeds, with_scores=False)
    dt = time.perf_counter() - t0
    its = np.array([r.iterations for r in runs])
    qps = nq / dt
    log(f"ppr: {nq} queries in {dt:.3f} s -> {qps:.1f} q/s, iterations mean {its.mean():.2f}")
    # the same batch with its tile sweeps replayed from captured chains of
    # four (GOPS_LAUNCH_GRAPH) instead of queued kernel by kernel
    from graphops_amd import _lib as L
    dg.ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_GRAPH)
    try:
        ga.ppr_batch_run(dg, cfg, off[:65], seeds[:off[64]], with_scores=False)
        t0 = time.perf_counter()
        ga.ppr_batch_run(dg, cfg, off, seeds, with_scores=False)
        gqps = nq / (time.perf_counter() - t0)
    finally:
        dg.ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_AUTO)
    log(f"ppr graph replay: {gqps:.1f} q/s")
    k = plain_sample
    t0 = time.perf_counter()
    _, pruns = ga.ppr_batch_run(dg, cfg, off[:k + 1], seeds[:off[k]], with_scores=False, tile=-1)
    pdt = time.perf_counter() - t0
    plain_qps = k / pdt
    log(f"ppr plain: {k} queries in {pdt:.3f} s -> {plain_qps:.2f} q/s")
    # accuracy on a sample: batched scores against the oracle's f64 run
    sample = [0, 1, 2, 3]
    sc, sruns = ga.ppr_batch_run(dg, cfg, off[:5], seeds[:off[4]], with_scores=True)
    refs = refs_of(oracle, e, seeds, off, sample)
    l1 = [float(np.abs(sc[j] - refs[j]).sum()) for j in sample]
    mx = [float(np.abs(sc[j] - refs[j]).max()) for j in sample]
    top = [top_agreement(sc[j], refs[j]) for j in sample]
    del sc
    alt = f64_phase_schedule(ga, dg, cfg, off, seeds, sample, refs)
    # CPU: the oracle's push form of personalized_pagerank_run, one thread
    t0 = time.perf_counter()
    for j in range(cpu_queries):
        p = np.zeros(e.n)
        p[seeds[off[j]:off[j + 1]]] = 1.0
        oracle.personalized_pagerank_run(e.offsets, e.targets, p, 0.85, 200, 1e-8)
    cdt = time.perf_counter() - t0
    n, m = e.n, e.m
    # byte model per query-sweep: 8n + (4m + 8n)/K; roofline q/s at 8 TB/s
    per_qsweep = 8 * n + (4 * m + 8 * n) / tile
    roof_qps = HBM_PEAK_GBS * 1e9 / (its.mean() * per_qsweep)
    return {
        "workload": f"rmat{PPR_GRAPH['scale']}-unit ppr {nq} queries |S|~U{{1..64}} "
                    "degree-proportional, d0.85 tol1e-8",
        "n": n, "m": m, "queries": nq, "queries_per_s": round(qps, 2),
        "ms_per_batch": round(dt * 1e3, 2), "tile": tile,
        "launch": "queued", "graph_replay_queries_per_s": round(gqps, 2),
        "tile_ms_per_query_sweep": {str(w): round(c, 5) for w, c in costs.items()},
        "iterations_mean": round(float(its.mean()), 3), "iterations_max": int(its.max()),
        "converged": int(sum(r.converged for r in runs)),
        "plain_queries_per_s": round(plain_qps, 3), "plain_sample_queries": k,
        "plain_iterations_mean": round(float(np.mean([r.iterations for r in pruns])), 3),
        "ratio_vs_plain": round(qps / plain_qps, 2),
        "accuracy": {"sample_queries": len(sample), "l1_vs_f64_oracle_max": max(l1),
                     "max_abs": max(mx), "top100_agreement_min": min(top)},
        "f64_phase_schedule": alt,
        "roofline": {"bound": "hbm", "queries_per_s_at_peak": round(roof_qps, 1),
                     "frac": round(qps / roof_qps, 4),
                     "bytes_per_query_sweep": round(per_qsweep),
                     "traffic_per_tile_sweep": measured_traffic("rmat22-unit ppr tile64", "tile")
                     if tile == 64 else None},
        "cpu_baseline": {"value": round(cpu_queries / cdt, 4), "unit": "queries/s", "cores": 1,
                         "kind": "port",
                         "sample": f"first {cpu_queries} queries by the oracle's "
                                   f"personalized_pagerank_run (f64, push), {cdt:.1f} s"},
    }


def refs_of(oracle, e, seeds, off, sample):
    """The oracle's f64 scores of the sample queries (d 0.85, tol 1e-8)."""
    in_off, in_src, _ = oracle.transpose(e.offsets, e.targets)
    deg = np.diff(e.offsets.astype(np.int64)).astype(np.float64)
    out = []
    for j in sample:
        p = np.zeros(e.n)
        p[seeds[off[j]:off[j + 1]]] += 1.0
        out.append(oracle.pull_run(2, in_off, in_src, deg, p=p, damping=0.85, max_iterations=200,
                                   tolerance=1e-8, threads=16).scores)
    return out


def f64_phase_schedule(ga, dg, cfg, off, seeds, sample, refs):
    """The same batch with a leading f64 phase (GOPS_OPT_TILE_F32_BELOW =
    1e4: f64 anchor sweeps from admission until the residual is within 1e4
    of the tolerance), the tiles' schedule before the carried f32 rounding
    made it unnecessary, with the L1 it reaches."""
    from graphops_amd import _lib as L

    nq = len(off) - 1
    ctx = dg.ctx
    ctx.set_option(L.OPT_TILE_F32_BELOW, 1e4)
    try:
        ga.ppr_batch_run(dg, cfg, off[:65], seeds[:off[64]], with_scores=False)  # warm-up
        t0 = time.perf_counter()
        _, runs = ga.ppr_batch_run(dg, cfg, off, seeds, with_scores=False)
        dt = time.perf_counter() - t0
        n_s = len(sample)
        sc, _ = ga.ppr_batch_run(dg, cfg, off[:n_s + 1], seeds[:off[n_s]], with_scores=True)
        l1 = max(float(np.abs(sc[i] - refs[i]).sum()) for i in range(n_s))
    finally:
        ctx.set_option(L.OPT_TILE_F32_BELOW, 0.0)
    qps = nq / dt
    log(f"ppr f64-phase schedule: {qps:.1f} q/s (L1 {l1:.2e})")
    return {"schedule": "f64 sweeps from admission while the residual > 1e4 * tol",
            "queries_per_s": round(qps, 2),
            "iterations_mean": round(float(np.mean([r.iterations for r in runs])), 3),
            "converged": int(sum(r.converged for r in runs)),
            "l1_vs_f64_oracle_max": l1}


def serve_once(ga, e, dg, off, seeds, damp, qps_offered, nq, max_slots, sweeps, topk, seed):
    """One serving run: nq Poisson arrivals at qps_offered, each response the
    query's top-k (node, score) selected on the device.  Latency =
    completion (poll) time - scheduled arrival time."""
    rng = np.random.default_rng(seed)
    arrive = np.cumsum(rng.exponential(1.0 / qps_offered, size=nq))
    srv = ga.Server(dg, tile=0, max_slots=max_slots, topk=topk)
    st = srv.stats()
    lat = np.zeros(nq)
    its = np.zeros(nq)
    conv = 0
    by_ticket = {}
    i = done = 0
    t0 = time.perf_counter()
    while done < nq:
        now = time.perf_counter() - t0
        while i < nq and arrive[i] <= now:
            t = srv.submit(ga.PageRankConfig(float(damp[i]), 1000, 1e-6), seeds[off[i]:off[i + 1]])
            by_ticket[t] = i
            i += 1
        s = srv.stats()
        if s["queued"] or s["active"]:
            srv.step(sweeps)
        elif i < nq:
            time.sleep(min(1e-4, max(0.0, arrive[i] - now)))
            continue
        while True:
            r = srv.poll(with_scores=False)
            if r is None:
                break
            j = by_ticket.pop(r[0])
            lat[j] = time.perf_counter() - t0 - arrive[j]
            its[j] = r[1].iterations
            conv += r[1].converged
            done += 1
    span = time.perf_counter() - t0
    srv.close()
    out = {"offered_qps": qps_offered, "queries": nq, "queries_per_s": round(nq / span, 2),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
           "iterations_mean": round(float(its.mean()), 2), "converged": int(conv),
           "slots": st["slots"]}
    log(f"serving at {qps_offered:.0f} QPS offered: {nq} queries in {span:.3f} s "
        f"({out['queries_per_s']} q/s), p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms")
    return out


def serving_leg(ga, generators, e, dg, nq=1024, loads=(5000.0, 10000.0, 50000.0),
                max_slots=512, sweeps=2, topk=100):
    """Serving loop (gops_server_*): Poisson arrivals, |S| log-uniform
    1..256, d uniform over {0.5, 0.85, 0.95}, tol 1e-6; continuous batching
    (admission into free slots between device steps); each response is the
    query's top-100, selected on the device.  Runs at the offered loads
    given (all above capacity: the throughput there is the capacity), then at
    half the measured capacity, where p50/p99 are service time rather than
    queue wait."""
    rng = np.random.default_rng(17)
    off, seeds = generators.seed_sets(e, nq, 1, 256, log_uniform=True, seed=23)
    damp = rng.choice([0.5, 0.85, 0.95], size=nq)
    runs = [serve_once(ga, e, dg, off, seeds, damp, q, nq, max_slots, sweeps, topk, 29 + k)
            for k, q in enumerate(loads)]
    capacity = max(r["queries_per_s"] for r in runs)
    low = round(0.5 * capacity, -1)
    n_low = min(nq, max(64, int(2.0 * low)))  # about two seconds of arrivals
    runs.append(serve_once(ga, e, dg, off, seeds, damp, low, n_low, max_slots, sweeps, topk, 41))
    return {"workload": f"rmat{PPR_GRAPH['scale']}-unit ppr serving, Poisson arrivals, "
                        "|S| log-uniform 1..256, d in {0.5,0.85,0.95}, tol1e-6",
            "capacity_queries_per_s": capacity, "loads": runs,
            "sweeps_per_step": sweeps, "response": f"top-{topk}"}


def topk_leg(ga, dg, scores, ks=(100, 1000), reps=20):
    """top_k (topk.rs:17-42) of the result of the last run, which stays on the
    device: gops_result_top_k selects it there and only k pairs cross PCIe
    (checked against the oracle's top_k of the same scores); the oracle's
    top_k on one core, over scores already on the host, is the CPU baseline."""
    import oracle

    out = {"n": dg.n}
    for k in ks:
        got = ga.result_top_k(dg, k)
        assert got == oracle.top_k(scores, k), f"top_k mismatch at k={k}"
        t0 = time.perf_counter()
        for _ in range(reps):
            ga.result_top_k(dg, k)
        out[f"device_ms_k{k}"] = round((time.perf_counter() - t0) / reps * 1e3, 3)
        log(f"top_k k={k}: {out[f'device_ms_k{k}']} ms on the device")
    t0 = time.perf_counter()
    oracle.top_k(scores, ks[0])
    cdt = time.perf_counter() - t0
    out["pcie_bytes"] = {"device_select": 12 * ks[0], "download_then_select": 8 * dg.n}
    out["cpu_baseline"] = {"value": round(cdt * 1e3, 2), "unit": "ms", "cores": 1, "kind": "port",
                           "sample": f"oracle top_k, k={ks[0]}, of the same {dg.n} scores"}
    return out


def build_leg(ga, e, storage):
    """Time gops_graph_create on the device build and on the host build (the
    same arrays): adapter lists in host memory -> graph resident on the GPU."""
    out = {}
    for where, reps in (("device", 3), ("host", 1)):
        cx = ga.Context(0, graph_build=where)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            g = ga.DeviceGraph.from_edges(e, weight_storage=storage, ctx=cx)
            ts.append(time.perf_counter() - t0)
            g.close()
        out[f"{where}_s"] = round(min(ts), 4)
        cx.close()
        log(f"graph build on the {where}: {min(ts):.3f} s (best of {reps})")
    out["speedup"] = round(out["host_s"] / out["device_s"], 2)
    out["device_gedges_per_s"] = round(e.m / out["device_s"] / 1e9, 3)
    out["host_threads"] = os.cpu_count() if not os.environ.get("OMP_NUM_THREADS") \
        else int(os.environ["OMP_NUM_THREADS"])
    return out


def centrality_leg(ga, e, dg, iters=20):
    """Katz, eigenvector and HITS on the resident graph (edge weights unused):
    iters iterations each at tolerance 0 (no early stop), results left on the
    device; the oracle's single-core Katz push sweep as the CPU baseline."""
    import ctypes

    import oracle
    from graphops_amd import _lib as L

    n, m = e.n, e.m
    lib, c, h = L.lib(), dg.ctx.handle, dg.handle
    info = L.RunInfoC()
    alpha = 0.85 / int(np.bincount(e.targets, minlength=n).max())

    def katz(k):
        cfg = L.KatzConfigC(alpha, 1.0, k, 0.0)
        rc = lib.gops_katz_centrality_run(c, h, ctypes.byref(cfg), None, ctypes.byref(info))
        assert rc == 0 and info.iterations == k

    def eig(k):
        rc = lib.gops_eigenvector_centrality_run(c, h, k, 0.0, None, ctypes.byref(info))
        assert rc == 0 and info.iterations == k

    def hits(k):
        rc = lib.gops_hits(c, h, k, 0.0, None, None, ctypes.byref(info))
        assert rc == 0 and info.iterations == k

    def timed(fn):
        fn(2)
        t0 = time.perf_counter()
        fn(iters)
        return (time.perf_counter() - t0) / iters

    out = {"iterations": iters, "katz_alpha": alpha}
    t0 = time.perf_counter()
    eig(1)  # builds the out-edge CSR once (and its blocked layout from 2^21 nodes)
    out["out_edge_csr_build_s"] = round(time.perf_counter() - t0, 4)
    # algorithmic bytes per iteration: 4 B of col per edge; per node the row
    # pointer, the gathered vector once and the vectors the epilogues and the
    # normalisations read and write
    for name, fn, sweeps, alg in (("katz", katz, 1, 4 * m + 36 * n),
                                  ("eigenvector", eig, 1, 4 * m + 52 * n),
                                  ("hits", hits, 2, 8 * m + 88 * n)):
        dt = timed(fn)
        out[name] = {"ms_per_iteration": round(dt * 1e3, 3),
                     "gteps": round(sweeps * m / dt / 1e9, 2),
                     "alg_bytes_per_iteration": alg,
                     "hbm_frac": round(alg / dt / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"{name}: {dt * 1e3:.3f} ms per iteration, {sweeps * m / dt / 1e9:.1f} GTEPS")
        # the default (blocked from 2^21 nodes) against the pull
        out[name]["sweep"] = "blocked" if dg.ctx.last_stats()["sweeps_blocked"] else "pull"
        dg.ctx.set_option(L.OPT_CENTRALITY_SWEEP, L.CENTRALITY_SWEEP_PULL)
        try:
            dp = timed(fn)
        finally:
            dg.ctx.set_option(L.OPT_CENTRALITY_SWEEP, L.CENTRALITY_SWEEP_AUTO)
        out[name]["pull_ms_per_iteration"] = round(dp * 1e3, 3)
        log(f"{name} as pulls: {dp * 1e3:.3f} ms per iteration")
    t0 = time.perf_counter()
    oracle.katz_run(e.offsets, e.targets, alpha, 1.0, 1, 0.0)
    cdt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": round(m / cdt / 1e9, 4), "unit": "GTEPS", "cores": 1,
                           "kind": "port", "sample": "one Katz push sweep of the oracle"}
    log(f"oracle Katz sweep on one core: {cdt:.2f} s")
    return out


def ppr_legs(args, ga, generators):
    t0 = time.perf_counter()
    e = generators.rmat(PPR_GRAPH["scale"], 16, 0.57, 0.19, 0.19, PPR_GRAPH["seed"])
    ctx = ga.Context(0, precision="mixed")
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    log(f"ppr graph n={e.n} m={e.m} resident in {time.perf_counter() - t0:.1f} s")
    out = {"ppr_batch": ppr_batch_leg(ga, generators, e, dg, nq=args.ppr_queries)}
    if not args.no_serving:
        out["serving"] = serving_leg(ga, generators, e, dg, nq=args.serve_queries)
    dg.close()
    return out


def copy_probe(nbytes, reps=5):
    """reps device copies of nbytes (torch's copy of a contiguous f32
    tensor): the known-size launch that scales the byte counters."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes // 4 * 4, "reps": reps}), flush=True)


def gather_probe(dg, reps=5):
    """gops_gather_probe: a kernel that makes only the f32 sweep's gathers on
    this graph (every edge's source value in stored order, the col stream, a
    sum); best of reps, HIP events, ms."""
    import ctypes

    from graphops_amd import _lib as L

    ms = ctypes.c_double()
    rc = L.lib().gops_gather_probe(dg.ctx.handle, dg.handle, reps, ctypes.byref(ms))
    assert rc == 0, rc
    return ms.value


def measured_traffic(workload, delta_sweep):
    """HBM bytes per f32 delta sweep of this workload and sweep kind from the
    committed counter summaries (FETCH_SIZE + WRITE_SIZE scaled by a copy of
    known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            records = json.load(f)
    except (OSError, ValueError):
        return None
    for t in records if isinstance(records, list) else [records]:
        if t.get("workload") == workload and t.get("delta_sweep", "pull") == delta_sweep:
            return t.get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-sweeps", type=int, default=3)
    ap.add_argument("--no-ppr", action="store_true", help="skip the batched PPR and serving legs")
    ap.add_argument("--ppr-queries", type=int, default=1024)
    ap.add_argument("--serve-queries", type=int, default=1024)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--ppr-only", action="store_true", help="only the PPR and serving legs")
    ap.add_argument("--no-build", action="store_true", help="skip the graph build timing leg")
    ap.add_argument("--build-only", action="store_true", help="only the graph build timing leg")
    ap.add_argument("--no-centrality", action="store_true", help="skip the Katz/eigenvector/HITS leg")
    ap.add_argument("--centrality-only", action="store_true",
                    help="only the Katz/eigenvector/HITS leg")
    ap.add_argument("--topk-only", action="store_true", help="only the top_k leg")
    ap.add_argument("--no-topk", action="store_true", help="skip the top_k leg")
    ap.add_argument("--no-gather-probe", action="store_true",
                    help="skip the gather-only probe of the sweep's access pattern")
    ap.add_argument("--delta-sweep", default="auto", choices=["auto", "pull", "blocked"],
                    help="how f32 delta sweeps read the graph (GOPS_OPT_DELTA_SWEEP)")
    ap.add_argument("--anchor-sweep", default="auto", choices=["auto", "pull"],
                    help="how f64 anchor sweeps read the graph with blocked delta sweeps "
                         "(GOPS_OPT_ANCHOR_SWEEP)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-c1c2", action="store_true", help="skip the C1 and C2 legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "graphops_amd runs on one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is not implemented)"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import graphops_amd as ga
    from graphops_amd import generators

    if args.ppr_only:
        print(json.dumps(ppr_legs(args, ga, generators)), flush=True)
        return
    c = CONFIGS[args.config]
    t0 = time.perf_counter()
    e = generators.rmat(c["scale"], 16, 0.57, 0.19, 0.19, c["seed"], c["weights"], c["p0"],
                        c["p1"], c["rounding"])
    log(f"generated n={e.n} m={e.m} in {time.perf_counter() - t0:.1f} s")
    if args.build_only:
        print(json.dumps({"workload": c["workload"], "n": e.n, "m": e.m,
                          "graph_build": build_leg(ga, e, c["storage"])}), flush=True)
        return
    if args.topk_only:
        cx = ga.Context(0)
        dg = ga.DeviceGraph.from_edges(e, weight_storage=c["storage"], ctx=cx)
        full = ga.pagerank_weighted_run(dg, ga.PageRankConfig(0.85, c["max_it"], c["tol"]),
                                        ctx=cx).scores
        print(json.dumps({"workload": c["workload"], "n": e.n,
                          "top_k": topk_leg(ga, dg, full)}), flush=True)
        return
    if args.centrality_only:
        dg = ga.DeviceGraph.from_edges(e, weight_storage=c["storage"],
                                       ctx=ga.Context(0, timing=False))
        print(json.dumps({"workload": f"rmat{c['scale']} unweighted", "n": e.n, "m": e.m,
                          "centrality": centrality_leg(ga, e, dg)}), flush=True)
        return
    ctx = ga.Context(0, precision="mixed", timing=True)
    from graphops_amd import _lib as L

    ctx.set_option(L.OPT_DELTA_SWEEP, {"auto": L.DELTA_SWEEP_AUTO, "pull": L.DELTA_SWEEP_PULL,
                                       "blocked": L.DELTA_SWEEP_BLOCKED}[args.delta_sweep])
    ctx.set_option(L.OPT_ANCHOR_SWEEP, {"auto": L.ANCHOR_SWEEP_AUTO,
                                        "pull": L.ANCHOR_SWEEP_PULL}[args.anchor_sweep])
    t0 = time.perf_counter()
    dg = ga.DeviceGraph.from_edges(e, weight_storage=c["storage"], ctx=ctx)
    log(f"graph resident on the GPU in {time.perf_counter() - t0:.1f} s")
    cfg = ga.PageRankConfig(0.85, c["max_it"], c["tol"])

    for i in range(args.warmup):
        r = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
        log(f"warmup {i}: {r.iterations} sweeps, diff {r.diff_l1:.3e}")
    # the timed runs take the library's default launch path; per-sweep kernel
    # times come from separate runs with every sweep timed, and the replay of
    # captured sweep chains (GOPS_LAUNCH_GRAPH) is timed beside them
    ctx.set_option(L.OPT_TIMING, 0)
    sweeps = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
        sweeps += r.iterations
    wall = time.perf_counter() - t0
    launch = "graph" if r.stats["graph_replay"] else "queued"
    ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_GRAPH)
    ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)  # captures the chain
    t0 = time.perf_counter()
    for i in range(2):
        ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
    graph_wall = (time.perf_counter() - t0) / 2
    ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_AUTO)
    ctx.set_option(L.OPT_TIMING, 1)
    f32_ms = f64_ms = 0.0
    n32 = n64 = 0
    qsweeps = 0
    t0 = time.perf_counter()
    for i in range(2):
        r = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
        st = r.stats
        qsweeps += r.iterations
        f32_ms += st["fp32_sweep_ms"]
        f64_ms += st["fp64_sweep_ms"]
        n32 += st["sweeps_fp32"]
        n64 += st["sweeps_fp64"]
    qwall = time.perf_counter() - t0
    anchors_blocked = st["sweeps_blocked"] == st["sweeps_fp32"] + st["sweeps_fp64"] > 0
    pull_anchor_us = None
    if anchors_blocked and st["sweeps_fp64"]:  # the same run with pull anchors, for comparison
        ctx.set_option(L.OPT_ANCHOR_SWEEP, L.ANCHOR_SWEEP_PULL)
        rq = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=False)
        ctx.set_option(L.OPT_ANCHOR_SWEEP, L.ANCHOR_SWEEP_AUTO)
        pull_anchor_us = rq.stats["fp64_sweep_ms"] / max(rq.stats["sweeps_fp64"], 1) * 1e3
    log(f"timed {args.steps} runs ({launch}): {wall:.3f} s, {sweeps} sweeps; graph replay "
        f"{graph_wall * 1e3:.2f} ms per run; every sweep timed {qwall / 2 * 1e3:.2f} ms per run")

    # the same runs with every result downloaded (n f64 over PCIe)
    t0 = time.perf_counter()
    pcie_runs = 2
    for _ in range(pcie_runs):
        rp = ga._run("gops_pagerank_weighted_run", dg, cfg, True, with_scores=True)
    pcie_wall = (time.perf_counter() - t0) / pcie_runs
    log(f"with the result downloaded: {pcie_wall * 1e3:.2f} ms per run")
    n, m = e.n, e.m
    value = m * sweeps / wall / 1e9
    alg_bytes = (4 + c["w_bytes"]) * m + 16 * n + 4
    avg_f32 = f32_ms / max(n32, 1) / 1e3
    avg_f64 = f64_ms / max(n64, 1) / 1e3
    avg_sweep = (f32_ms + f64_ms) / max(n32 + n64, 1) / 1e3
    # the dominant kernel: the f32 delta sweep (HIP events around its launches)
    achieved = alg_bytes / avg_f32 / 1e9 if n32 else None
    wname = "Bf16W" if c["w_bytes"] == 2 else "float"
    # which kind of f32 delta sweep actually ran, from the library's own count
    blocked = st["sweeps_blocked"] > 0
    sweep_name = (f"gops::pb_scatter_kernel<gops::{wname}> + gops::pb_accumulate_kernel + "
                  "gops::pb_combine_kernel (blocked f32 delta sweep)" if blocked
                  else f"gops::sweep_kernel<gops::{wname}> (f32 delta sweep)")
    out = {
        "metric": "PageRank GTEPS/iter (% HBM roofline) + batched PPR queries/s @ L1<=1e-6",
        "value": round(value, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic R-MAT (own counter-RNG generator)",
        "config": {"workload": c["workload"], "n": n, "m": m, "sweeps_per_run": r.iterations,
                   "f32_sweeps_per_run": st["sweeps_fp32"], "f64_sweeps_per_run": st["sweeps_fp64"],
                   "converged": r.converged, "diff_l1": r.diff_l1,
                   "precision": f"{'bf16' if c['w_bytes'] == 2 else 'f32'} weights, f32 delta "
                                "sweeps with f64 (pull) or fixed-point int64 (blocked) row sums, "
                                "f64 anchor sweeps (blocked: f64 run sums added as 128-bit "
                                "fixed point), f64 iterate, residual and dangling sums",
                   "delta_sweep": "blocked" if blocked else "pull",
                   "anchor_sweep": "blocked" if anchors_blocked else "pull",
                   "f64_pull_anchor_us": round(pull_anchor_us, 2) if pull_anchor_us else None,
                   "sweep_us": round(avg_sweep * 1e6, 2),
                   "f32_sweep_us": round(avg_f32 * 1e6, 2), "f64_sweep_us": round(avg_f64 * 1e6, 2),
                   "gteps_f32_sweep": round(m / avg_f32 / 1e9, 2) if n32 else None,
                   "launch": launch,
                   "graph_replay_ms_per_run": round(graph_wall * 1e3, 3),
                   "timed_kernels_ms_per_run": round(qwall / 2 * 1e3, 3)},
        "gather_bound": None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": measured_traffic(c["workload"], "blocked" if blocked else "pull"),
                     "kernel": sweep_name, "delta_sweep": "blocked" if blocked else "pull",
                     "alg_bytes_per_launch": alg_bytes},
        "pcie_inclusive": {"what": "whole runs with the n f64 scores downloaded to the host",
                           "ms_per_run": round(pcie_wall * 1e3, 3),
                           "gteps": round(m * rp.iterations / pcie_wall / 1e9, 3)},
    }
    try:
        out["roofline"]["measured_copy_peak"] = copy_peak()
    except Exception as exc:  # no HIP runtime: the nominal peak alone
        out["roofline"]["measured_copy_peak"] = {"error": str(exc)[:80]}
    if not args.no_gather_probe:
        pm = gather_probe(dg)
        out["gather_bound"] = {
            "probe": "gops_gather_probe: the pull sweep's gathers alone (source values in "
                     "stored order, the col stream, a sum)",
            "probe_ms": round(pm, 3), "gathers_per_s": round(m / pm / 1e6, 1),
            "f32_sweep_ms": round(avg_f32 * 1e3, 3),
            "f32_sweep": "blocked" if blocked else "pull",
            # the probe bounds the pull; a blocked sweep does not gather
            "frac": round(pm / (avg_f32 * 1e3), 3) if n32 and not blocked else None}
        log(f"gather probe: {pm:.3f} ms for {m} gathers; f32 sweep {avg_f32 * 1e3:.3f} ms")
    if not args.no_accuracy:
        scores = ga.pagerank_weighted_run(dg, cfg, ctx=ctx).scores
        t0 = time.perf_counter()
        ref, exact, ref_s = oracle_reference(e, c["tol"], c["max_it"])
        log(f"oracle runs: f64 {ref.iterations} sweeps, long double {exact.iterations} sweeps, "
            f"{time.perf_counter() - t0:.1f} s")
        diff = np.abs(scores - ref.scores)
        dx = np.abs(scores - exact.scores)
        out["accuracy"] = {"l1_vs_f64_oracle": float(diff.sum()), "max_abs": float(diff.max()),
                           "top100_agreement": top_agreement(scores, ref.scores),
                           "oracle_iterations": ref.iterations, "gpu_iterations": r.iterations,
                           "gpu_diff_l1": r.diff_l1,
                           "exact_iteration": {
                               "what": "the same iteration in extended precision (oracle, long "
                                       "double)",
                               "iterations": exact.iterations, "diff_l1": exact.diff_l1,
                               "l1_gpu_vs_exact": float(dx.sum()), "max_abs": float(dx.max()),
                               "l1_f64_oracle_vs_exact": float(np.abs(ref.scores -
                                                                      exact.scores).sum()),
                               "f64_oracle_mass_drift": float(ref.scores.sum() - 1.0)}}
        out["best_cpu"] = {"value": round(m * ref.iterations / ref_s / 1e9, 4), "unit": "GTEPS",
                           "cores": machine_info()["omp_threads"], "kind": "port",
                           "sample": f"the oracle's OpenMP pull (f64), the whole run "
                                     f"({ref.iterations} sweeps, {ref_s:.1f} s)"}
        del ref, exact
    out["cpu_baseline"] = cpu_baseline(e, c, args.cpu_sweeps) if args.cpu_sweeps > 0 else None
    if not args.no_c1c2:
        out["c1"] = c1_leg(ga, generators)
        log(f"C1: {out['c1']['gpu_ms_per_run']} ms per run on the GPU")
        out["c2"] = c2_leg(ga, generators)
        log(f"C2: {out['c2']['gteps']} GTEPS, f32 sweep {out['c2']['f32_sweep_us']} us")
    if not args.no_build:
        out["graph_build"] = build_leg(ga, e, c["storage"])
    if not args.no_centrality:
        out["centrality"] = centrality_leg(ga, e, dg)
    if not args.no_topk:  # top_k of a weighted run's result, which stays on the device
        full = ga.pagerank_weighted_run(dg, cfg, ctx=ctx).scores
        out["top_k"] = topk_leg(ga, dg, full)
        del full
    if not args.no_ppr:
        dg.close()
        del e
        out.update(ppr_legs(args, ga, generators))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
