"""Write tests/golden/golden.npz: independent golden vectors for the oracle.

Sources of truth are third-party or closed form, never the oracle itself:
  * BFS levels: networkx.single_source_shortest_path_length on a MultiDiGraph
  * SSSP distances: networkx.single_source_dijkstra_path_length (parallel
    edges keep their minimum weight)
  * PageRank: scipy.sparse power iteration in float64 with testPageRank.cu's
    stopping rule (stop after `max` updates or when sum |delta| < 1e-7)

Graphs: the reference's data files (tests/golden/data, copied verbatim from
the reference's data/ directory), a 128x128 grid, small seeded RMAT and
uniform graphs from the generator specification in oracle/__init__.py.

Run: python tests/golden/make_golden.py   (takes about a minute)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import networkx as nx  # noqa: E402
import scipy.sparse as sp  # noqa: E402

import oracle  # noqa: E402  (loader + generator spec only)

INF = 0x7FFFFFFF


def nx_graph(n, rp, col, w=None):
    G = nx.MultiDiGraph()
    G.add_nodes_from(range(n))
    src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    for i, (u, v) in enumerate(zip(src.tolist(), col.tolist())):
        G.add_edge(u, v, weight=int(w[i]) if w is not None else 1)
    return G


def nx_bfs(n, rp, col, s):
    G = nx_graph(n, rp, col)
    out = np.full(n, INF, dtype=np.int32)
    for v, d in nx.single_source_shortest_path_length(G, s).items():
        out[v] = d
    return out


def nx_sssp(n, rp, col, w, s):
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    for u, v, x in zip(src.tolist(), col.tolist(), np.asarray(w).tolist()):
        if G.has_edge(u, v):
            G[u][v]["weight"] = min(G[u][v]["weight"], x)
        else:
            G.add_edge(u, v, weight=x)
    out = np.full(n, INF, dtype=np.int32)
    for v, d in nx.single_source_dijkstra_path_length(G, s).items():
        out[v] = d
    return out


def scipy_pagerank(n, rp, col, max_iters, d=0.85, eps=1e-7):
    deg = np.diff(rp.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    P = sp.csr_matrix((1.0 / deg[src], (col.astype(np.int64), src)), shape=(n, n))  # duplicates summed
    r = np.full(n, 1.0 / n)
    it = 0
    while it < max_iters:
        nr = d * (P @ r) + (1 - d) / n
        err = np.abs(nr - r).sum()
        r = nr
        if err < eps:
            break
        it += 1
    return r


def main():
    out = {}
    data = os.path.join(HERE, "data")
    for f in sorted(os.listdir(data)):
        n, rp, col, w = oracle.load_edge_list(os.path.join(data, f))
        out[f"{f}/row_ptr"] = rp
        out[f"{f}/col"] = col
        srcs = [0] if n < 3 else [0, n // 2, n - 1]
        for s in srcs:
            out[f"{f}/bfs/{s}"] = nx_bfs(n, rp, col, s)
            if w is not None:
                out[f"{f}/sssp/{s}"] = nx_sssp(n, rp, col, w, s)
        if col.size:
            out[f"{f}/pr20"] = scipy_pagerank(n, rp, col, 20)
            out[f"{f}/pr1000"] = scipy_pagerank(n, rp, col, 1000)
        else:
            out[f"{f}/pr20"] = scipy_pagerank(n, rp, np.zeros(0, np.uint32), 20)
    # 128 x 128 grid (the CPU-path BFS configuration)
    n, rp, col = oracle.gen_grid(128, 128)
    out["grid128/bfs/0"] = nx_bfs(n, rp, col, 0)
    # small RMAT and uniform graphs from the generator spec
    n, rp, col = oracle.gen_rmat(10, 16, seed=7)
    out["rmat10/row_ptr"], out["rmat10/col"] = rp, col
    for s in (0, 1, 37):
        out[f"rmat10/bfs/{s}"] = nx_bfs(n, rp, col, s)
    out["rmat10/pr20"] = scipy_pagerank(n, rp, col, 20)
    n, rp, col = oracle.gen_uniform(2000, 16000, seed=3)
    w = oracle.weights_i32(col.size, 1, 64, 3)
    out["unif2k/row_ptr"], out["unif2k/col"], out["unif2k/w"] = rp, col, w
    for s in (0, 999):
        out[f"unif2k/sssp/{s}"] = nx_sssp(n, rp, col, w, s)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
