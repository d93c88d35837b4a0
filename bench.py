"""Benchmark of the frontier engine on one MI355X.

Default workload (`--workload pagerank`): PageRank on RMAT scale 20
(1,048,576 vertices, edge factor 16, a/b/c/d = .57/.19/.19/.05, self-loops
dropped, seed 1), damping 0.85, exactly 20 iterations, fp32 ranks with fp32
accumulation -- the configuration the headline metric is quoted on.  One
step = one full 20-iteration PageRank over the HBM-resident graph.  The line
also carries `secondary` measurements of the other targets (skip them with
--no-secondary):
  bfs_rmat22      BFS on RMAT-22 (8 seeded sources), with the cost of the
                  first call (renumbered twin built) and the graph-order time
  grid128         BFS on the 128 x 128 grid from vertex 0: the oracle on one
                  host core (the configuration's CPU path) and the GPU
  batched_256     256-source BFS on RMAT-22 by plan replay vs 256 plain calls
  multigraph_64   64 RMAT graphs (scales 12..18) in one captured plan vs 64
                  plain calls
  adjoint_rmat24  weighted PageRank forward + VJP on RMAT-24 vs forward-only
                  with central finite differences

Prints one JSON line.  `value` is edges/s = E * 20 * steps / time.
`roofline` covers the dominant kernel (the fused pull iteration):
algorithmic bytes per launch 8E + 12V over its hipEvent-timed average launch
duration.  `cpu_baseline` is the oracle's serial fp64 PageRank (restating
testPageRank.cu:57-104) on one host core over the same graph.

Other workloads print their own line with the same keys:
  bfs      BFS on RMAT-22 from seeded sources (GTEPS, bytes 8E_r+16V_r+4V)
  sssp     SSSP on the uniform 2^22 x 2^25 graph, int32 w in [1,64], source 0
  batched  256-source BFS on RMAT-22 and the 64-graph RMAT batch (scales
           12..18) as one captured persistent launch, vs the plain engine
  adjoint  weighted PageRank forward + VJP on RMAT-24 vs forward-only with
           central finite differences
  ingest   RMAT-22 edge-list text (edges in random order) parsed on the
           device into the out-CSR and in-CSR (edges/s, text resident in HBM)
  egress   RMAT-22 BFS levels written as "%d\n" text (BFS.cu's output file),
           formatted on the device (lines/s, file in the temp directory)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS (BFS/SSSP), PageRank edges/s + %HBM peak; batched/adjoint vs plain"
INF = 0x7FFFFFFF


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic(kernel_key):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get(kernel_key)
    except Exception:
        return None


def base_line(args, workload_cfg, value, unit, ms_per_step, dtype, higher=True):
    return {
        "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded generators)", "config": workload_cfg,
    }


def roofline(bytes_per_launch, avg_launch_s, key, kernel):
    achieved = bytes_per_launch / avg_launch_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(key), "kernel": kernel,
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_s * 1e6}


_CPU = {"build": "prebuilt liboracle.so (-O3, baseline x86-64)"}


def cpu_setup():
    """The CPU baseline's oracle: oracle.c built here with -O3 -march=native
    (the host that runs this), falling back to the prebuilt portable build."""
    import tempfile
    import oracle
    if _CPU.get("native_tried"):
        return
    _CPU["native_tried"] = True
    try:
        path = oracle.build_native(tempfile.mkdtemp(prefix="olive_oracle_"))
        oracle.use_library(path)
        _CPU["build"] = "gcc -O3 -march=native on this host"
    except Exception as e:  # no compiler: keep the portable build
        log(f"native oracle build failed ({e}); using the prebuilt one")


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def timed_oracle(fn, budget_s, max_runs=5):
    """Median wall time of up to max_runs calls within budget_s, with this
    process pinned to one core (the oracle is serial)."""
    cpu_setup()
    allowed = os.sched_getaffinity(0)
    core = min(allowed)
    os.sched_setaffinity(0, {core})
    times = []
    try:
        t_start = time.perf_counter()
        while len(times) < max_runs and (time.perf_counter() - t_start) < budget_s:
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, allowed)
    _CPU["core"] = core
    return float(np.median(times)), len(times)


def cpu_line(value, unit, sample):
    """cpu_baseline object: one pinned core, the host described."""
    return {"value": value, "unit": unit, "cores": 1, "kind": "port", "sample": sample,
            "build": _CPU["build"], "pinned_core": _CPU.get("core"), "host_cpu": host_cpu(),
            "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def sources_of(csr, k, seed):
    """k sources among vertices with out-edges (Graph500 convention)."""
    nz = np.flatnonzero(csr.outdeg())
    rng = np.random.default_rng(seed)
    return rng.choice(nz, size=k, replace=k > nz.size).astype(np.uint32)


# ------------------------------------------------------------------ PageRank
def bench_pagerank(args, olv):
    import oracle
    iters = 20
    log("generating RMAT-20 ...")
    csr = olv.gen_rmat(20, 16, seed=1)
    g = olv.Graph(csr)
    n, m = csr.n, csr.m
    out = olv.DeviceBuffer(n, np.float32)
    L = olv._lib.lib()
    st = olv.Stats()
    # the first call also builds the in-degree relabelled in-CSR and its
    # bucket plan (once per graph)
    t_first = time.perf_counter()
    olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, 0, ctypes.byref(st)))
    L.olv_synchronize()
    t_first = time.perf_counter() - t_first
    for _ in range(args.warmup):
        olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, 0, ctypes.byref(st)))
    L.olv_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, 0, ctypes.byref(st)))
    L.olv_synchronize()
    dt = time.perf_counter() - t0
    # per-launch kernel time for the roofline: separate runs with hipEvents
    # around every iteration launch (the events themselves add gaps, so these
    # runs are not the timed ones)
    kms, kl = 0.0, 0
    for _ in range(max(1, min(args.steps, 5))):
        olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, olv._lib.OLV_PROFILE, ctypes.byref(st)))
        kms += st.kernel_ms
        kl += st.kernel_launches
    # accuracy of the last run against the fp64 oracle (not timed)
    ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, iters, epsilon=0.0)
    got = out.numpy().copy()
    err = float(np.max(np.abs(got.astype(np.float64) - ref) / ref))
    # the per-row pull in the graph's own vertex order and over the
    # renumbered CSR: per-launch times for comparison (the same ranks up to
    # the summation order)
    gms, gl, rms, rl = 0.0, 0, 0.0, 0
    if not args.no_modes:
        for _ in range(3):
            olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c,
                                              olv._lib.OLV_PROFILE | olv._lib.OLV_GRAPH_ORDER, ctypes.byref(st)))
            gms += st.kernel_ms
            gl += st.kernel_launches
        g_order = out.numpy().copy()
        for _ in range(3):
            olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c,
                                              olv._lib.OLV_PROFILE | olv._lib.OLV_PR_ROW_PULL, ctypes.byref(st)))
            rms += st.kernel_ms
            rl += st.kernel_launches
    cfg = {"workload": "pagerank_rmat20", "vertices": n, "edges": m, "iterations": iters, "damping": 0.85,
           "rmat": [0.57, 0.19, 0.19, 0.05], "seed": 1, "max_rel_err_vs_fp64": err,
           "first_call_ms": t_first * 1e3}
    if gl:
        cfg["graph_order_launch_us"] = gms / gl * 1e3
        cfg["row_pull_launch_us"] = rms / rl * 1e3
        cfg["row_pull_bit_identical_to_graph_order"] = bool(np.array_equal(out.numpy().view(np.uint32),
                                                                           g_order.view(np.uint32)))
        cfg["max_rel_diff_vs_row_pull"] = float(np.max(np.abs(got.astype(np.float64) - out.numpy()) /
                                                       out.numpy().astype(np.float64)))
    res = base_line(args, cfg, m * iters * args.steps / dt, "edges/s", dt / args.steps * 1e3, "f32")
    # one iteration = the segment pass and the combine (hipEvents around both)
    res["roofline"] = roofline(8 * m + 12 * n, kms / kl / 1e3, "pagerank_rmat20",
                               "blk_segment_kernel + blk_combine_kernel (one iteration)")
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        t, runs = timed_oracle(lambda: oracle.pagerank(csr.row_ptr, csr.col, iters, epsilon=0.0), 20.0)
        res["cpu_baseline"] = cpu_line(m * iters / t, "edges/s",
                                       f"full workload (RMAT-20, {iters} iterations), serial fp64 restatement of "
                                       f"testPageRank.cu:57-104, median of {runs} runs")
    if not args.no_secondary:
        del g, out
        sec = {}
        r22, csr22, g22, _ = bfs_rmat22(olv, args, cpu=not args.no_cpu_baseline)
        sec["bfs_rmat22"] = r22
        sec["grid128"] = grid128(olv, args)
        sec.update(batched_legs(olv, args, csr22, g22, sssp=False))
        del g22, csr22
        sec["api_bfs"] = api_bfs_legs()
        sec["adjoint_rmat24"] = adjoint_leg(olv, args)
        res["secondary"] = sec
    return res


# ----------------------------------------------------------------------- BFS
def bfs_rmat22(olv, args, cpu=True, nsrc=8):
    import oracle
    log("generating RMAT-22 ...")
    csr = olv.gen_rmat(22, 16, seed=1)
    g = olv.Graph(csr)
    srcs = sources_of(csr, nsrc, 22)
    out = olv.DeviceBuffer(csr.n, np.int32)
    L = olv._lib.lib()
    # first call on the graph: work buffers and the in-degree renumbered twin
    # are built inside it (wall clock, then one steady call for comparison)
    L.olv_synchronize()
    t0 = time.perf_counter()
    olv._lib.check(L.olv_bfs(g.h, int(srcs[0]), out.c, 0, None))
    L.olv_synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    olv._lib.check(L.olv_bfs(g.h, int(srcs[0]), out.c, 0, None))
    L.olv_synchronize()
    t_steady = time.perf_counter() - t0
    for s in srcs[:2]:
        for _ in range(max(1, args.warmup)):
            g.bfs(int(s), out=out)
    per = []
    for s in srcs:
        _, st = g.bfs(int(s), out=out)
        per.append((st.elapsed_ms / 1e3, st.edges_traversed, st.vertices_reached, st.iterations,
                    st.bottom_up_levels))
    t = np.array([p[0] for p in per])
    er = np.array([p[1] for p in per], np.float64)
    vr = np.array([p[2] for p in per], np.float64)
    teps = er / t
    gteps_h = len(teps) / np.sum(1.0 / teps) / 1e9  # harmonic mean (Graph500)
    bytes_ = 8 * er + 16 * vr + 4 * csr.n
    frac = float(np.mean(bytes_ / t / 1e9 / HBM_PEAK_GBS))
    check = oracle.bfs(csr.row_ptr, csr.col, int(srcs[-1]))
    ok = bool(np.array_equal(out.numpy(), check))
    r = {"gteps_harmonic": gteps_h, "hbm_frac_mean": frac, "ms_mean": float(t.mean() * 1e3),
         "edges": csr.m, "sources": len(srcs), "levels": [p[3] for p in per], "bottom_up_levels": [p[4] for p in per],
         "bit_exact_last_source": ok, "first_call_ms": t_first * 1e3, "steady_call_ms": t_steady * 1e3,
         "twin_build_ms": (t_first - t_steady) * 1e3,
         "twin_build_in_traversals": (t_first - t_steady) / max(float(t.mean()), 1e-9)}
    # the same sources through the other level drivers (same results)
    modes = {"host_loop": olv.OLV_HOST_LOOP, "persistent": olv.OLV_PERSISTENT, "push_only": olv.OLV_NO_DIRECTION_OPT,
             "graph_order": olv.OLV_GRAPH_ORDER}
    r["modes_ms_mean"] = {"pipeline": r["ms_mean"]}
    for name, fl in ({} if args.no_modes else modes).items():
        ts = []
        for s in srcs:
            g.bfs(int(s), out=out, flags=fl)
            ts.append(g.bfs(int(s), out=out, flags=fl)[1].elapsed_ms)
        r["modes_ms_mean"][name] = float(np.mean(ts))
    if "graph_order" in r["modes_ms_mean"]:
        r["graph_order_ms_mean"] = r["modes_ms_mean"]["graph_order"]
    if cpu:
        tc, runs = timed_oracle(lambda: oracle.bfs(csr.row_ptr, csr.col, int(srcs[0])), 20.0)
        r["cpu_gteps"] = float(er[0] / tc / 1e9)
        r["cpu_baseline"] = cpu_line(r["cpu_gteps"], "GTEPS", f"one source of the same graph, serial restatement "
                                     f"of testBFS.cu:56-79, median of {runs}")
    return r, csr, g, srcs


def bench_bfs(args, olv):
    r, csr, g, srcs = bfs_rmat22(olv, args, cpu=not args.no_cpu_baseline, nsrc=args.sources)
    # roofline of the persistent BFS kernel (one launch per traversal)
    cfg = {"workload": "bfs_rmat22", "vertices": csr.n, "edges": csr.m, "sources": r["sources"],
           "rmat": [0.57, 0.19, 0.19, 0.05], "seed": 1, "bit_exact": r["bit_exact_last_source"]}
    res = base_line(args, cfg, r["gteps_harmonic"], "GTEPS", r["ms_mean"], "int32")
    res["roofline"] = {"bound": "hbm", "achieved": r["hbm_frac_mean"] * HBM_PEAK_GBS, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": r["hbm_frac_mean"], "traffic": load_traffic("bfs_rmat22"),
                       "kernel": "bfs_expand_kernel + bfs_compact_kernel (whole traversal)",
                       "bytes_model": "8*E_r + 16*V_r + 4*V"}
    if "cpu_baseline" in r:
        res["cpu_baseline"] = r.pop("cpu_baseline")
    res["detail"] = r
    return res


# ---------------------------------------------------------------------- SSSP
def bench_sssp(args, olv):
    import oracle
    log("generating uniform 2^22 x 2^25 ...")
    csr = olv.with_weights_i32(olv.gen_uniform(1 << 22, 1 << 25, seed=1), 1, 64, seed=1)
    g = olv.Graph(csr, transpose=False)
    out = olv.DeviceBuffer(csr.n, np.int32)
    for _ in range(args.warmup):
        g.sssp(0, out=out)
    ts = []
    for _ in range(args.steps):
        _, st = g.sssp(0, out=out)
        ts.append(st.elapsed_ms / 1e3)
    t = float(np.median(ts))
    er, vr = st.edges_traversed, st.vertices_reached
    ref = oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, 0)
    ok = bool(np.array_equal(out.numpy(), ref))
    cfg = {"workload": "sssp_uniform_4M_33M", "vertices": csr.n, "edges": csr.m, "weights": [1, 64], "source": 0,
           "bit_exact": ok, "rounds": st.iterations, "edges_relaxed": st.edges_inspected}
    res = base_line(args, cfg, er / t / 1e9, "GTEPS", t * 1e3, "int32")
    b = 12 * er + 16 * vr + 4 * csr.n
    res["roofline"] = {"bound": "hbm", "achieved": b / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": b / t / 1e9 / HBM_PEAK_GBS, "traffic": load_traffic("sssp_uniform"),
                       "kernel": "ssspf_round_kernel (whole run: one launch per round)",
                       "bytes_model": "12*E_r + 16*V_r + 4*V"}
    if not args.no_cpu_baseline:
        tc, runs = timed_oracle(lambda: oracle.sssp(csr.row_ptr, csr.col, csr.weights, 0), 30.0, 2)
        td, druns = timed_oracle(lambda: oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, 0), 30.0, 2)
        res["cpu_baseline"] = cpu_line(er / tc / 1e9, "GTEPS", f"full workload, serial frontier relaxation "
                                       f"(SSSP.cu:146-163), median of {runs}")
        res["cpu_baseline"]["dijkstra_gteps"] = er / td / 1e9
    return res


# ------------------------------------------------------------------- batched
def batched_legs(olv, args, csr, g, sssp=True):
    """256-source BFS on RMAT-22 by plan replay vs the plain engine, the
    64-graph batch in one captured plan vs 64 plain calls, and (sssp) 16-source
    SSSP on the uniform graph batched vs plain.  Returns a dict of legs."""
    import oracle
    import tempfile
    L = olv._lib.lib()
    reps = max(3, args.steps // 4)
    srcs = sources_of(csr, 256, 4)
    # launch shapes and thresholds from the autotuner (measured on this graph,
    # or taken from the tuning cache under its signature)
    t0 = time.perf_counter()
    tune = g.autotune_ex(olv._lib.OLV_TUNE_BFS | olv._lib.OLV_TUNE_BATCHED,
                         os.path.join(tempfile.gettempdir(), "olive_amd_tuning.jsonl"))
    t_tune = time.perf_counter() - t0
    out = olv.DeviceBuffer(csr.n * 256, np.int32)
    plan = ctypes.c_void_p()
    t0 = time.perf_counter()
    olv._lib.check(L.olv_plan_bfs_batched(g.h, srcs.ctypes.data, 256, out.c, 0, ctypes.byref(plan)))
    L.olv_synchronize()
    t_capture = time.perf_counter() - t0
    st = olv.Stats()
    for _ in range(args.warmup):
        olv._lib.check(L.olv_plan_launch(plan, ctypes.byref(st)))
    tb, tw = [], []
    for _ in range(reps):
        L.olv_synchronize()
        t0 = time.perf_counter()
        olv._lib.check(L.olv_plan_launch(plan, ctypes.byref(st)))
        L.olv_synchronize()
        tw.append(time.perf_counter() - t0)
        tb.append(st.elapsed_ms / 1e3)
    t_batched, t_batched_wall = float(np.median(tb)), float(np.median(tw))
    got = out.numpy().reshape(256, csr.n)
    rows = (0, 97, 255)
    exact = all(np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(srcs[i]))) for i in rows)
    del got
    # plain engine, one source at a time (same sources): wall clock over 256
    # C-ABI calls without statistics, and the sum of their device times
    # (hipEvents from reset to last level: the plain engine's best case)
    one = olv.DeviceBuffer(csr.n, np.int32)
    er_tot, t_plain_dev = 0, 0.0
    for s in srcs:
        _, s1 = g.bfs(int(s), out=one)
        er_tot += s1.edges_traversed
        t_plain_dev += s1.elapsed_ms / 1e3
    L.olv_synchronize()
    t0 = time.perf_counter()
    for s in srcs:
        olv._lib.check(L.olv_bfs(g.h, int(s), one.c, 0, None))
    L.olv_synchronize()
    t_plain = time.perf_counter() - t0
    tuning = tune.as_dict()
    tuning["autotune_wall_ms"] = t_tune * 1e3
    tuning["signature"] = g.signature()
    L.olv_plan_destroy(plan)
    del out, one
    legs = {"batched_256": {
        "graph": "rmat22", "sources": 256, "batched_ms": t_batched * 1e3, "batched_wall_ms": t_batched_wall * 1e3,
        "capture_ms": t_capture * 1e3, "plain_256_wall_ms": t_plain * 1e3, "plain_256_device_ms": t_plain_dev * 1e3,
        "speedup_wall": t_plain / t_batched_wall, "speedup_device": t_plain_dev / t_batched,
        "gteps": er_tot / t_batched / 1e9, "edges_traversed": int(er_tot), "bit_exact_sampled_rows": list(rows),
        "bit_exact": bool(exact), "tuning": tuning}}
    # 64 RMAT graphs, scales uniform in [12, 18], one captured launch
    log("generating the 64-graph batch ...")
    csrs = [olv.gen_rmat(olv.batch_scale(1, i, 12, 18), 16, seed=olv.batch_seed(1, i)) for i in range(64)]
    gs = [olv.Graph(c) for c in csrs]
    bsrc = np.array([int(np.flatnonzero(c.outdeg())[0]) for c in csrs], np.uint32)
    arr = (ctypes.c_void_p * 64)(*[x.h.value for x in gs])
    uni = ctypes.c_void_p()
    offs = np.zeros(65, np.uint32)
    olv._lib.check(L.olv_graph_union(arr, 64, ctypes.byref(uni), offs.ctypes.data))
    seeds = (offs[:64] + bsrc).astype(np.uint32)
    utune = olv._lib.Tuning()
    t0 = time.perf_counter()
    olv._lib.check(L.olv_autotune_ex(uni, olv._lib.OLV_TUNE_BFS,
                                     os.path.join(tempfile.gettempdir(), "olive_amd_tuning.jsonl").encode(),
                                     ctypes.byref(utune)))
    t_utune = time.perf_counter() - t0
    uout = olv.DeviceBuffer(int(offs[64]), np.int32)
    uplan = ctypes.c_void_p()
    olv._lib.check(L.olv_plan_bfs(uni, seeds.ctypes.data, 64, uout.c, 0, ctypes.byref(uplan)))
    for _ in range(args.warmup):
        olv._lib.check(L.olv_plan_launch(uplan, ctypes.byref(st)))
    tu, tuw = [], []
    for _ in range(reps):
        L.olv_synchronize()
        t0 = time.perf_counter()
        olv._lib.check(L.olv_plan_launch(uplan, ctypes.byref(st)))
        L.olv_synchronize()
        tuw.append(time.perf_counter() - t0)
        tu.append(st.elapsed_ms / 1e3)
    t_union, t_union_wall = float(np.median(tu)), float(np.median(tuw))
    uo = uout.numpy()
    uexact = all(np.array_equal(uo[offs[i]:offs[i + 1]], oracle.bfs(csrs[i].row_ptr, csrs[i].col, int(bsrc[i])))
                 for i in range(64))
    # plain engine, one graph at a time: outputs allocated and every graph
    # run once beforehand (work buffers, renumbered twin), then timed
    outs = [olv.DeviceBuffer(c.n, np.int32) for c in csrs]
    for x, s, o in zip(gs, bsrc, outs):
        x.bfs(int(s), out=o)
    L.olv_synchronize()
    t0 = time.perf_counter()
    for x, s, o in zip(gs, bsrc, outs):
        olv._lib.check(L.olv_bfs(x.h, int(s), o.c, 0, None))
    L.olv_synchronize()
    t_plain_multi = time.perf_counter() - t0
    L.olv_plan_destroy(uplan)
    L.olv_graph_destroy(uni)
    legs["multigraph_64"] = {
        "graphs": 64, "scales": [12, 18], "vertices": int(offs[64]), "edges": int(sum(c.m for c in csrs)),
        "plan_ms": t_union * 1e3, "plan_wall_ms": t_union_wall * 1e3, "plain_64_wall_ms": t_plain_multi * 1e3,
        "speedup_wall": t_plain_multi / t_union_wall, "bit_exact_all_graphs": bool(uexact),
        "tuning": dict(utune.as_dict(), autotune_wall_ms=t_utune * 1e3)}
    del gs, outs, uout
    if not sssp:
        return legs
    # 16-source SSSP on the uniform 2^22 x 2^25 graph, batched vs plain
    log("generating uniform 2^22 x 2^25 ...")
    ucsr = olv.with_weights_i32(olv.gen_uniform(1 << 22, 1 << 25, seed=1), 1, 64, seed=1)
    ug = olv.Graph(ucsr, transpose=False)
    ssrc = sources_of(ucsr, 16, 5)
    sout, sst = ug.sssp_batched(ssrc)
    ts = []
    for _ in range(reps):
        sout, sst = ug.sssp_batched(ssrc, out=sout)
        ts.append(sst.elapsed_ms / 1e3)
    t_sssp_b = float(np.median(ts))
    one = olv.DeviceBuffer(ucsr.n, np.int32)
    ug.sssp(int(ssrc[0]), out=one)
    t_sssp_p = 0.0
    sexact = True
    sb = sout.numpy().reshape(16, ucsr.n)
    for i, s_ in enumerate(ssrc):
        _, s1 = ug.sssp(int(s_), out=one)
        t_sssp_p += s1.elapsed_ms / 1e3
        if i in (0, 15):
            sexact = sexact and bool(np.array_equal(sb[i], one.numpy()))
    legs["sssp_batched_16"] = {"graph": "uniform_4M_33M", "sources": 16, "batched_ms": t_sssp_b * 1e3,
                               "plain_16_device_ms": t_sssp_p * 1e3, "speedup": t_sssp_p / t_sssp_b,
                               "batched_equals_plain": sexact}
    return legs


def bench_batched(args, olv):
    log("generating RMAT-22 ...")
    csr = olv.gen_rmat(22, 16, seed=1)
    g = olv.Graph(csr)
    legs = batched_legs(olv, args, csr, g, sssp=True)
    b = legs["batched_256"]
    cfg = {"workload": "batched_bfs_rmat22_256src_plus_64_graphs", "vertices": csr.n, "edges": csr.m,
           "sources": 256, "bit_exact_sampled": bool(b["bit_exact"] and legs["multigraph_64"]["bit_exact_all_graphs"]),
           "multigraph_vertices": legs["multigraph_64"]["vertices"],
           "multigraph_edges": legs["multigraph_64"]["edges"]}
    res = base_line(args, cfg, b["gteps"], "GTEPS", b["batched_ms"], "int32")
    res["detail"] = legs
    return res


# ------------------------------------------------------------ Olive C++ API
def api_bfs_legs():
    """The reference BFS program's level loop (edgeFilter / vertexFilter)
    through the Olive C++ API, beside olv_bfs on the same graph and sources:
    olive_amd/_apps/olive_api_bench (built from olive_amd/examples), run as
    a child process on RMAT-22 (8 sources) and the 128 x 128 grid."""
    import subprocess
    exe = os.path.join(ROOT, "olive_amd", "_apps", "olive_api_bench")
    out = {}
    for name, argv in (("rmat22", ["rmat", "22", "8"]), ("grid128", ["grid", "128", "128"])):
        if not os.path.exists(exe):
            out[name] = {"error": "olive_api_bench not built"}
            continue
        r = subprocess.run([exe] + argv, capture_output=True, text=True, timeout=300)
        try:
            out[name] = json.loads(r.stdout.strip().splitlines()[-1])
        except (ValueError, IndexError):
            out[name] = {"error": f"exit {r.returncode}", "stderr": r.stderr[-300:]}
    return out


# --------------------------------------------------------------------- grid
def grid128(olv, args):
    """BFS on the 128 x 128 directed grid from vertex 0 (255 levels): the
    oracle on one host core -- this configuration's CPU reference path --
    and the GPU engine, both over the same graph."""
    import oracle
    csr = olv.gen_grid(128, 128)
    g = olv.Graph(csr)
    out = olv.DeviceBuffer(csr.n, np.int32)
    for _ in range(max(1, args.warmup)):
        g.bfs(0, out=out)
    ts = []
    for _ in range(max(5, args.steps // 2)):
        _, st = g.bfs(0, out=out)
        ts.append(st.elapsed_ms / 1e3)
    t_gpu = float(np.median(ts))
    ref = oracle.bfs(csr.row_ptr, csr.col, 0)
    lv = out.numpy()
    closed = np.add.outer(np.arange(128), np.arange(128)).ravel()
    r = {"vertices": csr.n, "edges": csr.m, "levels": int(lv.max()) + 1, "gpu_ms": t_gpu * 1e3,
         "gpu_gteps": csr.m / t_gpu / 1e9, "bit_exact": bool(np.array_equal(lv, ref) and np.array_equal(lv, closed))}
    if not args.no_cpu_baseline:
        tc, runs = timed_oracle(lambda: oracle.bfs(csr.row_ptr, csr.col, 0), 5.0, max_runs=5)
        r["cpu_ms"] = tc * 1e3
        r["cpu_baseline"] = cpu_line(csr.m / tc / 1e9, "GTEPS", f"the whole workload, serial restatement of "
                                     f"testBFS.cu:56-79, median of {runs}")
    return r


# ------------------------------------------------------------------- adjoint
def adjoint_leg(olv, args, scale=24):
    """Weighted PageRank forward + VJP on RMAT-`scale` (tol 1e-6) against
    forward-only with central finite differences (2 forward solves per
    direction).  Repeated calls pass OLV_W_UNCHANGED (the weights are the
    same), so the weights are put in slot order once per solver."""
    log(f"generating weighted RMAT-{scale} ...")
    csr = olv.with_weights_f32_normalized(olv.gen_rmat(scale, 16, seed=1), seed=1)
    t = olv.gen_dirichlet(csr.n, seed=1)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    rng = np.random.default_rng(1)
    gbar = olv.DeviceBuffer.from_numpy(rng.standard_normal(csr.n).astype(np.float32))
    U = olv._lib.OLV_W_UNCHANGED
    t0 = time.perf_counter()
    r, st = g.pagerank_weighted(W, T, tol=1e-6)
    _, _, _, sa = g.pagerank_vjp(W, r, gbar, tol=1e-6)
    olv._lib.lib().olv_synchronize()
    t_first = time.perf_counter() - t0
    fwd, adj = [], []
    for _ in range(max(2, args.steps // 4)):
        r, st = g.pagerank_weighted(W, T, tol=1e-6, out=r, flags=U)
        fwd.append(st.elapsed_ms)
        _, _, _, sa = g.pagerank_vjp(W, r, gbar, tol=1e-6, flags=U)
        adj.append(sa.elapsed_ms)
    t_f, t_a = float(np.median(fwd)), float(np.median(adj))
    # the same calls re-permuting the weights every time (no reuse flag)
    r, st0 = g.pagerank_weighted(W, T, tol=1e-6, out=r)
    _, _, _, sa0 = g.pagerank_vjp(W, r, gbar, tol=1e-6)
    # one directional derivative by central differences = 2 forward solves
    t_fd_dir = 2 * t_f
    it_f = st.iterations
    d = {"vertices": csr.n, "edges": csr.m, "tol": 1e-6, "forward_ms": t_f, "vjp_ms": t_a,
         "forward_iterations": it_f, "vjp_iterations": sa.iterations,
         "forward_iteration_ms": t_f / max(it_f, 1),
         "forward_iteration_hbm_frac": (12 * csr.m + 12 * csr.n) / (t_f / max(it_f, 1) / 1e3) / 1e9 / HBM_PEAK_GBS,
         "forward_ms_repermuting_weights": st0.elapsed_ms, "vjp_ms_repermuting_weights": sa0.elapsed_ms,
         "first_calls_ms": t_first * 1e3, "fd_per_direction_ms": t_fd_dir,
         "fd_full_gradient_ms": t_fd_dir * (csr.m + csr.n),
         "speedup_full_gradient_vs_fd": t_fd_dir * (csr.m + csr.n) / (t_f + t_a),
         "fd_directions_for_5x": (t_f + t_a) * 5 / t_fd_dir}
    if not args.no_modes:
        go = olv.OLV_GRAPH_ORDER
        g.pagerank_weighted(W, T, tol=1e-6, out=r, flags=go)
        g.pagerank_vjp(W, r, gbar, tol=1e-6, flags=go)
        d["graph_order_forward_ms"] = g.pagerank_weighted(W, T, tol=1e-6, out=r, flags=go | U)[1].elapsed_ms
        d["graph_order_vjp_ms"] = g.pagerank_vjp(W, r, gbar, tol=1e-6, flags=go)[3].elapsed_ms
    return d


def bench_adjoint(args, olv, scale=24):
    d = adjoint_leg(olv, args, scale)
    cfg = {"workload": f"pagerank_fwd_adjoint_rmat{scale}", "vertices": d["vertices"], "edges": d["edges"],
           "tol": 1e-6, "fwd_iterations": d["forward_iterations"], "adj_iterations": d["vjp_iterations"]}
    t_f, t_a = d["forward_ms"], d["vjp_ms"]
    res = base_line(args, cfg, d["edges"] * (d["forward_iterations"] + d["vjp_iterations"] + 1) / ((t_f + t_a) / 1e3),
                    "edges/s", t_f + t_a, "f32")
    res["detail"] = d
    return res


# -------------------------------------------------------------------- ingest
def bench_ingest(args, olv):
    """One step = edge-list text already in HBM -> resident graph with both
    CSRs (olv_graph_from_edge_list_text + OLV_BUILD_TRANSPOSE)."""
    import oracle
    L = olv._lib.lib()
    log("generating RMAT-22 and its edge-list text ...")
    csr = olv.gen_rmat(22, 16, seed=1)
    src = np.repeat(np.arange(csr.n, dtype=np.uint32), np.diff(csr.row_ptr.astype(np.int64)))
    perm = np.random.default_rng(1).permutation(csr.m)
    s_, d_ = src[perm], csr.col[perm]
    text = olv.format_edge_list(csr.n, s_, d_, comment="RMAT-22, edges in random order")
    dtext = olv.DeviceBuffer(len(text), np.uint8)
    olv._lib.check(L.olv_memcpy_h2d(dtext.ptr, text, len(text)))
    flags = olv._lib.OLV_BUILD_TRANSPOSE
    st = olv.Stats()

    def ingest(ptr, fl, stats):
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        olv._lib.check(L.olv_graph_from_edge_list_text(ctypes.byref(h), ptr, len(text), fl, ctypes.byref(stats)))
        dt = time.perf_counter() - t0
        return h, dt

    for _ in range(args.warmup):
        h, _ = ingest(dtext.ptr, flags, st)
        L.olv_graph_destroy(h)
    times = []
    for _ in range(args.steps):
        h, dt = ingest(dtext.ptr, flags, st)
        times.append(dt)
        if len(times) < args.steps:
            L.olv_graph_destroy(h)
    t_step = float(np.median(times))
    # parity: the device CSR equals the stable sort by source, the device
    # in-CSR equals the host transpose of the generator's CSR
    g = olv.Graph._adopt(h, csr=None)
    order = np.argsort(s_, kind="stable")
    exact = bool(np.array_equal(g.csr.row_ptr, csr.row_ptr) and np.array_equal(g.csr.col, d_[order]))
    del g
    # PCIe-inclusive: the same text from pageable host memory
    h, t_host = ingest(text, flags | olv._lib.OLV_PTR_HOST, st)
    L.olv_graph_destroy(h)
    # dominant kernel (line parsing), hipEvent-timed in separate profiled runs
    pk = []
    for _ in range(3):
        h, _ = ingest(dtext.ptr, flags | olv._lib.OLV_PROFILE, st)
        pk.append(st.kernel_ms / 1e3)
        nlines = st.iterations
        L.olv_graph_destroy(h)
    k_s = float(np.median(pk))
    cfg = {"workload": "ingest_rmat22_edge_list_text", "vertices": csr.n, "edges": csr.m, "text_bytes": len(text),
           "lines": int(nlines), "transpose": True, "bit_exact": exact}
    res = base_line(args, cfg, csr.m / t_step, "edges/s", t_step * 1e3, "u8")
    # parse kernel: text + line starts (8) + content index (4) + src/dst/w out (12) per line
    res["roofline"] = roofline(len(text) + 12 * nlines + 12 * csr.m, k_s, "ingest_rmat22_parse", "edge_parse_kernel")
    res["detail"] = {"ms_device_resident": t_step * 1e3, "ms_from_host_text": t_host * 1e3,
                     "text_GBps_device_resident": len(text) / t_step / 1e9, "parse_kernel_ms": k_s * 1e3}
    if not args.no_cpu_baseline:
        # the oracle's C loader (csrGraph.h:106-171 restated + stable counting sort), same text, one core
        t_cpu, runs = timed_oracle(lambda: oracle.load_edge_list_text(text), 30.0, max_runs=2)
        res["cpu_baseline"] = cpu_line(csr.m / t_cpu, "edges/s",
                                       f"the whole RMAT-22 text, oracle.c orc_edge_list_load, median of {runs}")
    return res


# -------------------------------------------------------------------- egress
def bench_egress(args, olv):
    """One step = the 4,194,304 BFS levels of RMAT-22 (HBM-resident) written
    as one "%d\n" line each to a file, as BFS.cu:153-155 does."""
    import tempfile
    import oracle
    log("generating RMAT-22 and its BFS levels ...")
    csr = olv.gen_rmat(22, 16, seed=1)
    g = olv.Graph(csr)
    src = int(sources_of(csr, 1, 4)[0])
    levels, _ = g.bfs(src)
    tmpd = tempfile.mkdtemp()
    path = os.path.join(tmpd, "BFS.txt")
    st = olv.Stats()
    for _ in range(args.warmup):
        olv.write_text(path, levels, "%d")
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        olv.write_text(path, levels, "%d", stats=st)
        times.append(time.perf_counter() - t0)
    t_step = float(np.median(times))
    text_bytes = st.edges_inspected
    lv = levels.numpy()
    with open(path, "rb") as fh:
        exact = fh.read() == "".join(f"{x}\n" for x in lv.tolist()).encode()
    pk = []
    for _ in range(3):
        olv.write_text(path, levels, "%d", stats=st, profile=True)
        pk.append(st.kernel_ms / max(st.kernel_launches, 1) / 1e3)
    k_s = float(np.median(pk))
    cfg = {"workload": "egress_rmat22_bfs_levels_text", "values": csr.n, "text_bytes": int(text_bytes),
           "format": "%d\n", "bit_exact": bool(exact)}
    res = base_line(args, cfg, csr.n / t_step, "lines/s", t_step * 1e3, "int32")
    # write pass: value 4 + offset 4 in, the text out
    res["roofline"] = roofline(8 * csr.n + text_bytes, k_s, "egress_rmat22_levels", "text_write_kernel")
    res["detail"] = {"ms_per_file": t_step * 1e3, "write_kernel_ms": k_s * 1e3}
    if not args.no_cpu_baseline:
        cpath = os.path.join(tmpd, "BFS_cpu.txt")
        t_cpu, runs = timed_oracle(lambda: oracle.print_values(cpath, lv), 30.0, max_runs=3)
        with open(cpath, "rb") as a, open(path, "rb") as b:
            res["config"]["bit_exact"] = bool(exact and a.read() == b.read())
        os.unlink(cpath)
        res["cpu_baseline"] = cpu_line(csr.n / t_cpu, "lines/s", f"the same levels, oracle.c orc_print_values "
                                       f"(fprintf per vertex), median of {runs}")
    os.unlink(path)
    os.rmdir(tmpd)
    return res


def bench_copy(args, olv):
    """Device-to-device copy of a known size: the calibration kernel for the
    PMC byte counters (FETCH_SIZE / WRITE_SIZE) of the other workloads."""
    nbytes = 256 << 20
    a = olv.DeviceBuffer(nbytes // 4, np.uint32)
    b = olv.DeviceBuffer(nbytes // 4, np.uint32)
    a.fill(7)
    L = olv._lib.lib()
    for _ in range(args.warmup):
        olv._lib.check(L.olv_memcpy_d2d(b.ptr, a.ptr, nbytes))
    L.olv_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        olv._lib.check(L.olv_memcpy_d2d(b.ptr, a.ptr, nbytes))
    L.olv_synchronize()
    dt = time.perf_counter() - t0
    cfg = {"workload": "copy_256MiB", "bytes_per_copy": nbytes}
    res = base_line(args, cfg, 2 * nbytes * args.steps / dt / 1e9, "GB/s", dt / args.steps * 1e3, "u8")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pagerank",
                    choices=["pagerank", "bfs", "sssp", "batched", "adjoint", "ingest", "egress", "copy"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--sources", type=int, default=8, help="bfs: seeded sources timed (default 8)")
    ap.add_argument("--no-modes", action="store_true", help="bfs: skip the host-loop/persistent/push-only drivers; pagerank: skip the graph-order comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one GPU holding the whole graph; "
                                   "multi-GPU runs are not part of this build (use --gpus 1)"}))
        sys.exit(2)
    import olive_amd as olv
    olv.init(0)
    fn = {"pagerank": bench_pagerank, "bfs": bench_bfs, "sssp": bench_sssp, "batched": bench_batched,
          "adjoint": bench_adjoint, "ingest": bench_ingest, "egress": bench_egress,
          "copy": bench_copy}[args.workload]
    res = fn(args, olv)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
