"""CPU oracle for the frontier engine -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``olive_amd``) never calls it and has no CPU fallback.

Contents
  * ctypes bindings to ``liboracle.so`` (oracle.c): BFS (testBFS.cu:56-79),
    frontier SSSP (SSSP.cu:146-163 over oliver.h:56-113), Dijkstra, fp64
    PageRank (testPageRank.cu:57-104), weighted PageRank and its adjoint.
  * file loaders restating csrGraph.h:106-171 (edge list) and
    csrGraph.h:177-236 (DIMACS), with the stable sort by source that the
    reference comments out (csrGraph.h:149) and the weight column it drops
    (csrGraph.h:137-140).
  * numpy restatements of the seeded generators' specification, used to pin
    the product generators bit-for-bit on small sizes.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
INF = 0x7FFFFFFF

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH) or (
        os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "oracle.c"))
    ):
        subprocess.run(["make", "-s", "-B" if force else "-s", "liboracle.so"], cwd=HERE, check=True)
    return LIB_PATH


def build_native(out_dir: str) -> str:
    """Compile oracle.c with -O3 -march=native into out_dir (the CPU
    baseline: built on the host that times it).  Returns the library path."""
    subprocess.run(["make", "-s", "-B", "native", f"NATIVE_DIR={os.path.abspath(out_dir)}"], cwd=HERE, check=True)
    return os.path.join(os.path.abspath(out_dir), "liboracle.so")


def use_library(path: str) -> None:
    """Bind the oracle functions to the build at `path` (e.g. the native
    build from build_native) instead of the default liboracle.so."""
    global _lib
    _lib = None
    lib(path)


def lib(path: str | None = None):
    global _lib
    if _lib is None:
        path = path or LIB_PATH
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.orc_bfs.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, i32p]
        L.orc_sssp_frontier.argtypes = [ctypes.c_uint32, u32p, u32p, i32p, ctypes.c_uint32, i32p]
        L.orc_dijkstra.argtypes = [ctypes.c_uint32, u32p, u32p, i32p, ctypes.c_uint32, i32p]
        L.orc_pagerank.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, f64p, f64p]
        L.orc_pagerank_weighted.argtypes = [ctypes.c_uint32, u32p, u32p, f64p, f64p, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, f64p]
        L.orc_pagerank_adjoint.argtypes = [ctypes.c_uint32, u32p, u32p, f64p, f64p, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, f64p]
        i64p = ctypes.POINTER(ctypes.c_longlong)
        L.orc_edge_list_header.argtypes = [ctypes.c_char_p, ctypes.c_size_t, i64p, i64p]
        L.orc_edge_list_load.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                         u32p, u32p, i32p, ctypes.POINTER(ctypes.c_int)]
        L.orc_print_values.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        L.orc_bfs_verify.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, u32p, i32p, ctypes.c_int]
        L.orc_bfs_verify.restype = ctypes.c_int64
        for f in ("orc_bfs", "orc_sssp_frontier", "orc_dijkstra", "orc_pagerank",
                  "orc_pagerank_weighted", "orc_pagerank_adjoint", "orc_edge_list_header",
                  "orc_edge_list_load", "orc_print_values"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _csr(row_ptr, col):
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    c = np.ascontiguousarray(col, dtype=np.uint32)
    if c.size == 0:
        c = np.zeros(1, dtype=np.uint32)
    return rp, c


# ---------------------------------------------------------------- algorithms
def bfs(row_ptr, col, src):
    """Levels from ``src`` (testBFS.cu:56-79); INF = 0x7fffffff."""
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    out = np.empty(max(n, 1), dtype=np.int32)
    rc = lib().orc_bfs(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32), src, _p(out, ctypes.c_int32))
    if rc < 0:
        raise ValueError(f"orc_bfs failed ({rc})")
    return out[:n]


def bfs_verify(row_ptr, col, srcs, levels, threads=None):
    """Index of the first row of `levels` (rows x n, row r from srcs[r])
    that is not the BFS level array, or -1 when all are (orc_bfs_verify:
    level 0 only at the source, no edge skips a level, every reached vertex
    has a parent one level up)."""
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    s = np.ascontiguousarray(srcs, dtype=np.uint32)
    lv = np.ascontiguousarray(levels, dtype=np.int32).reshape(s.size, n)
    th = threads or min(16, len(os.sched_getaffinity(0)))
    rc = lib().orc_bfs_verify(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32), s.size,
                              _p(s, ctypes.c_uint32), _p(lv, ctypes.c_int32), th)
    if rc < -1:
        raise MemoryError("orc_bfs_verify: allocation failed")
    return int(rc)


def sssp(row_ptr, col, w, src):
    """Frontier relaxation distances (SSSP.cu:146-163); returns (dist, rounds)."""
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    ww = np.ascontiguousarray(w, dtype=np.int32)
    if ww.size == 0:
        ww = np.zeros(1, dtype=np.int32)
    out = np.empty(max(n, 1), dtype=np.int32)
    rc = lib().orc_sssp_frontier(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32),
                                 _p(ww, ctypes.c_int32), src, _p(out, ctypes.c_int32))
    if rc < -1:
        raise ValueError(f"orc_sssp_frontier failed ({rc})")
    return out[:n], rc


def dijkstra(row_ptr, col, w, src):
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    ww = np.ascontiguousarray(w, dtype=np.int32)
    if ww.size == 0:
        ww = np.zeros(1, dtype=np.int32)
    out = np.empty(max(n, 1), dtype=np.int32)
    rc = lib().orc_dijkstra(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32),
                            _p(ww, ctypes.c_int32), src, _p(out, ctypes.c_int32))
    if rc < 0:
        raise ValueError(f"orc_dijkstra failed ({rc})")
    return out[:n]


def pagerank(row_ptr, col, max_iters=20, damping=0.85, epsilon=1e-7):
    """fp64 PageRank (testPageRank.cu:57-104); returns (ranks, updates, err)."""
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    out = np.empty(n, dtype=np.float64)
    err = ctypes.c_double(0.0)
    rc = lib().orc_pagerank(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32), max_iters,
                            damping, epsilon, _p(out, ctypes.c_double), ctypes.byref(err))
    if rc < 0:
        raise ValueError(f"orc_pagerank failed ({rc})")
    return out, rc, err.value


def pagerank_weighted(row_ptr, col, w, t, damping=0.85, tol=1e-12, max_iters=10000, allow_cap=False):
    """Returns (ranks, iterations).  allow_cap: hitting max_iters is not an
    error (used to run an exact number of updates)."""
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    ww = np.ascontiguousarray(w, dtype=np.float64)
    if ww.size == 0:
        ww = np.zeros(1)
    tt = np.ascontiguousarray(t, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    rc = lib().orc_pagerank_weighted(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32),
                                     _p(ww, ctypes.c_double), _p(tt, ctypes.c_double), damping, tol,
                                     max_iters, _p(out, ctypes.c_double))
    if rc == -1 and allow_cap:
        return out, max_iters
    if rc < 0:
        raise ValueError(f"orc_pagerank_weighted did not converge ({rc})")
    return out, rc


def pagerank_adjoint(row_ptr, col, w, gbar, damping=0.85, tol=1e-14, max_iters=10000):
    rp, c = _csr(row_ptr, col)
    n = rp.size - 1
    ww = np.ascontiguousarray(w, dtype=np.float64)
    if ww.size == 0:
        ww = np.zeros(1)
    g = np.ascontiguousarray(gbar, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    rc = lib().orc_pagerank_adjoint(n, _p(rp, ctypes.c_uint32), _p(c, ctypes.c_uint32),
                                    _p(ww, ctypes.c_double), _p(g, ctypes.c_double), damping, tol,
                                    max_iters, _p(out, ctypes.c_double))
    if rc < 0:
        raise ValueError(f"orc_pagerank_adjoint did not converge ({rc})")
    return out, rc


def pagerank_vjp(row_ptr, col, w, r, gbar, damping=0.85, tol=1e-14):
    """fp64 VJP of r = d P^T r + (1-d) t for L = <gbar, r>: (grad_w, grad_t, lambda)."""
    lam, _ = pagerank_adjoint(row_ptr, col, w, gbar, damping, tol)
    rp = np.asarray(row_ptr, dtype=np.int64)
    src = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    grad_w = damping * np.asarray(r, np.float64)[src] * lam[np.asarray(col, np.int64)]
    return grad_w, (1.0 - damping) * lam, lam


# ------------------------------------------------------------------- loaders
def load_edge_list(path, with_weights=True):
    """Edge-list file -> (n, row_ptr, col, weights|None).

    Restates csrGraph.h:106-171: '#' lines skipped, first other line is
    ``n m``, then ``src dst [w]`` lines.  Unlike the reference the tuples are
    stable-sorted by source (the sort is commented out at csrGraph.h:149) and
    the optional third column is kept (dropped at csrGraph.h:139).
    """
    header = None
    rows = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            parts = s.split()
            if header is None:
                header = (int(parts[0]), int(parts[1]))
                continue
            rows.append([int(x) for x in parts])
    n, m = header
    rows = rows[:m]
    src = np.array([r[0] for r in rows], dtype=np.int64)
    dst = np.array([r[1] for r in rows], dtype=np.int64)
    has_w = with_weights and len(rows) > 0 and all(len(r) >= 3 for r in rows)
    w = np.array([r[2] for r in rows], dtype=np.int32) if has_w else None
    order = np.argsort(src, kind="stable")
    src, dst = src[order], dst[order]
    if w is not None:
        w = w[order]
    row_ptr = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr, dtype=np.uint64).astype(np.uint32)
    return n, row_ptr, dst.astype(np.uint32), w


def load_edge_list_text(text: bytes):
    """The C restatement of the edge-list loader (oracle.c
    orc_edge_list_load) over text in memory: (n, row_ptr, col, w|None)."""
    n, m = ctypes.c_longlong(), ctypes.c_longlong()
    if lib().orc_edge_list_header(text, len(text), ctypes.byref(n), ctypes.byref(m)) != 0:
        raise ValueError("no header line")
    n, m = n.value, m.value
    rp = np.empty(n + 1, np.uint32)
    col = np.empty(max(m, 1), np.uint32)
    w = np.empty(max(m, 1), np.int32)
    has_w = ctypes.c_int()
    rc = lib().orc_edge_list_load(text, len(text), n, m, _p(rp, ctypes.c_uint32), _p(col, ctypes.c_uint32),
                                  _p(w, ctypes.c_int32), ctypes.byref(has_w))
    if rc != 0:
        raise ValueError(f"orc_edge_list_load failed ({rc})")
    return n, rp, col[:m], (w[:m] if has_w.value else None)


def print_values(path, values):
    """The reference programs' per-vertex fprintf loop (oracle.c
    orc_print_values): "%d\n" for int32 values, "%f\n" for float64."""
    a = np.ascontiguousarray(values)
    if a.dtype not in (np.int32, np.float64):
        raise ValueError("int32 or float64 values")
    if lib().orc_print_values(str(path).encode(), a.ctypes.data, a.size, int(a.dtype == np.float64)) != 0:
        raise OSError(f"cannot write {path}")


def load_dimacs(path):
    """METIS/DIMACS adjacency file -> (n, row_ptr, col), csrGraph.h:177-236.

    Header ``n m`` (m undirected edges, 2m arcs), then one line per vertex
    listing 1-based neighbour ids.
    """
    with open(path) as fh:
        lines = [ln for ln in fh.read().split("\n")]
    body = [ln for ln in lines if not ln.lstrip().startswith("#")]
    it = iter(body)
    header = None
    for ln in it:
        if ln.strip():
            header = ln.split()
            break
    n, m = int(header[0]), int(header[1])
    adj = []
    for ln in it:
        if len(adj) == n:
            break
        adj.append([int(x) - 1 for x in ln.split()])
    while len(adj) < n:
        adj.append([])
    row_ptr = np.zeros(n + 1, dtype=np.uint32)
    row_ptr[1:] = np.cumsum([len(a) for a in adj])
    col = np.array([x for a in adj for x in a], dtype=np.uint32)
    assert col.size == 2 * m, (col.size, m)
    return n, row_ptr, col


# --------------------------------------------- generator spec (numpy, small)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _draw(seed, idx):
    with np.errstate(over="ignore"):
        return _mix64(_mix64(np.uint64(seed)) + np.asarray(idx, dtype=np.uint64) * _GOLD)


def _coo_to_csr(n, src, dst):
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    row_ptr = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(row_ptr, src.astype(np.int64) + 1, 1)
    return np.cumsum(row_ptr).astype(np.uint32), dst.astype(np.uint32)


def gen_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, drop_self_loops=True):
    """RMAT: per edge, per level (MSB first) one 32-bit uniform picks the
    quadrant (a: 00, b: 01, c: 10, d: 11); two levels per 64-bit draw (high
    word first)."""
    n = 1 << scale
    m = n * edge_factor
    words = (scale + 1) // 2
    thr = [min(int(np.floor(p * 4294967296.0)), 1 << 32) for p in (a, a + b, a + b + c)]
    i = np.arange(m, dtype=np.uint64)
    u = np.zeros(m, dtype=np.uint64)
    v = np.zeros(m, dtype=np.uint64)
    for wd in range(words):
        z = _draw(seed, i * np.uint64(words) + np.uint64(wd))
        for half in range(2):
            level = 2 * wd + half
            if level >= scale:
                break
            r = (z >> np.uint64(32)) if half == 0 else (z & np.uint64(0xFFFFFFFF))
            q = (r >= thr[0]).astype(np.uint64) + (r >= thr[1]) + (r >= thr[2])
            bit = np.uint64(scale - 1 - level)
            u |= (q >> np.uint64(1)) << bit
            v |= (q & np.uint64(1)) << bit
    if drop_self_loops:
        keep = u != v
        u, v = u[keep], v[keep]
    return (n,) + _coo_to_csr(n, u, v)


def gen_uniform(n, m, seed=1):
    z = _draw(seed, np.arange(m, dtype=np.uint64))
    with np.errstate(over="ignore"):
        u = ((z >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)
        v = ((z & np.uint64(0xFFFFFFFF)) * np.uint64(n)) >> np.uint64(32)
    return (n,) + _coo_to_csr(n, u, v)


def gen_grid(rows, cols):
    src, dst = [], []
    for r in range(rows):
        for c in range(cols):
            i = r * cols + c
            for (rr, cc) in ((r - 1, c), (r, c - 1), (r, c + 1), (r + 1, c)):
                if 0 <= rr < rows and 0 <= cc < cols:
                    src.append(i)
                    dst.append(rr * cols + cc)
    return (rows * cols,) + _coo_to_csr(rows * cols, np.array(src, np.uint64), np.array(dst, np.uint64))


def weights_i32(m, lo, hi, seed):
    z = _draw(np.uint64(seed) ^ np.uint64(0x5EED0001), np.arange(m, dtype=np.uint64))
    rng = np.uint64(hi - lo + 1)
    with np.errstate(over="ignore"):
        return (lo + (((z >> np.uint64(32)) * rng) >> np.uint64(32)).astype(np.int64)).astype(np.int32)


def weights_f32_normalized(row_ptr, seed):
    m = int(row_ptr[-1])
    z = _draw(np.uint64(seed) ^ np.uint64(0x5EED0002), np.arange(m, dtype=np.uint64))
    raw = ((z >> np.uint64(40)) + np.uint64(1)).astype(np.float64) / 16777216.0
    src = np.repeat(np.arange(row_ptr.size - 1), np.diff(row_ptr.astype(np.int64)))
    sums = np.zeros(row_ptr.size - 1)
    # sequential per-row sum in CSR order, like the product generator
    for v in range(row_ptr.size - 1):
        s = 0.0
        for k in range(int(row_ptr[v]), int(row_ptr[v + 1])):
            s += raw[k]
        sums[v] = s
    return (raw / sums[src]).astype(np.float32)
