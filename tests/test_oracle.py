"""Pin the CPU oracle before trusting it (runs without a GPU).

Each oracle function is checked against something other than itself:
golden vectors from networkx / scipy (tests/golden/golden.npz), closed forms
on the reference's data files, invariants, and dense linear algebra.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import DATA_FILES, GOLDEN, data_file

INF = oracle.INF
G = np.load(os.path.join(GOLDEN, "golden.npz"))


def load(name):
    n, rp, col, w = oracle.load_edge_list(data_file(name))
    return n, rp, col, w


# ------------------------------------------------------------------ loaders
def test_loader_matches_golden_csr():
    for f in DATA_FILES:
        n, rp, col, _ = load(f)
        assert np.array_equal(rp, G[f"{f}/row_ptr"]) and np.array_equal(col, G[f"{f}/col"])


def test_loader_ordered_graph_by_hand():
    # orderedGraph_9 edge list read off the file by hand
    n, rp, col, w = load("orderedGraph_9")
    assert n == 9 and w is None
    assert rp.tolist() == [0, 2, 5, 5, 6, 8, 9, 9, 11, 11]
    assert col.tolist() == [1, 3, 0, 2, 4, 4, 5, 7, 8, 6, 8]


def test_loader_sorts_disordered_input():
    # the reference loader corrupts this file (its sort is commented out);
    # sorted by source it is the same multigraph as orderedGraph_9
    _, rp1, c1, _ = load("orderedGraph_9")
    _, rp2, c2, _ = load("disorderedGraph_9")
    assert np.array_equal(rp1, rp2)
    for v in range(9):
        assert sorted(c1[rp1[v]:rp1[v + 1]]) == sorted(c2[rp2[v]:rp2[v + 1]])


def test_loader_keeps_weights():
    n, rp, col, w = load("gridGraph_15")
    src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    assert np.array_equal(w, (src + col).astype(np.int32))  # the file's weights are u+v


# --------------------------------------------------------------------- BFS
def test_bfs_closed_forms():
    n, rp, col, _ = load("gridGraph_15")  # 3 x 5 grid, id = 5*row + col
    lv = oracle.bfs(rp, col, 0)
    assert lv.tolist() == [(i // 5) + (i % 5) for i in range(15)]
    n, rp, col, _ = load("chainGraph_1K")
    assert oracle.bfs(rp, col, 0).tolist() == list(range(1000))
    for f in ("starGraph_1K", "acyclicGraph_100"):
        n, rp, col, _ = load(f)
        assert oracle.bfs(rp, col, 0).tolist() == [0] + [1] * (n - 1)
    n, rp, col, _ = load("isolatedGraph_1K")
    assert oracle.bfs(rp, col, 0).tolist() == [0] + [INF] * 999
    n, rp, col, _ = load("singletonGraph_1")
    assert oracle.bfs(rp, col, 0).tolist() == [0]
    for f in ("orderedGraph_9", "disorderedGraph_9"):
        n, rp, col, _ = load(f)
        assert oracle.bfs(rp, col, 0).tolist() == [0, 1, 2, 1, 2, 3, 4, 3, 4]


def test_bfs_matches_networkx_golden():
    keys = [k for k in G.files if "/bfs/" in k]
    assert len(keys) > 20
    for k in keys:
        name, _, s = k.split("/")
        if name == "grid128":
            _, rp, col = oracle.gen_grid(128, 128)
        elif name.startswith("rmat"):
            rp, col = G[f"{name}/row_ptr"], G[f"{name}/col"]
        else:
            _, rp, col, _ = load(name)
        assert np.array_equal(oracle.bfs(rp, col, int(s)), G[k]), k


def test_bfs_grid128_closed_form():
    _, rp, col = oracle.gen_grid(128, 128)
    lv = oracle.bfs(rp, col, 0)
    i = np.arange(128 * 128)
    assert np.array_equal(lv, (i // 128 + i % 128).astype(np.int32))
    assert lv.max() == 254


def bfs_invariants(rp, col, src, lv):
    n = rp.size - 1
    s = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    reached = lv[s] != INF
    # lvl[v] <= lvl[u] + 1 along every edge out of a reached vertex
    assert np.all(lv[col[reached]].astype(np.int64) <= lv[s[reached]].astype(np.int64) + 1)
    # every reached v != src has an in-neighbour one level up
    has_parent = np.zeros(n, bool)
    ok = reached & (lv[col] == lv[s] + 1)
    has_parent[col[ok]] = True
    r = (lv != INF)
    r[src] = False
    assert np.all(has_parent[r])


def test_bfs_invariants_rmat():
    _, rp, col = oracle.gen_rmat(12, 16, seed=5)
    for s in (0, 11, 4000):
        bfs_invariants(rp, col, s, oracle.bfs(rp, col, s))


def test_bfs_plants_mistake():
    # a planted off-by-one in a copy of the oracle's output fails the invariants
    _, rp, col = oracle.gen_rmat(10, 16, seed=7)
    lv = oracle.bfs(rp, col, 0).copy()
    v = int(np.flatnonzero((lv != INF) & (lv > 1))[0])
    lv[v] += 1
    with pytest.raises(AssertionError):
        bfs_invariants(rp, col, 0, lv)


def test_bfs_verify_rows():
    """The multi-row C verifier accepts the oracle's rows (and the numpy
    invariants agree) and flags each planted mistake at its row."""
    _, rp, col = oracle.gen_rmat(11, 16, seed=3)
    outdeg = np.diff(rp.astype(np.int64))
    srcs = np.flatnonzero(outdeg)[[0, 5, 50, 500, 900]].astype(np.uint32)
    rows = np.stack([oracle.bfs(rp, col, int(s)) for s in srcs])
    for s, r in zip(srcs, rows):
        bfs_invariants(rp, col, int(s), r)
    for th in (1, 3, 8):
        assert oracle.bfs_verify(rp, col, srcs, rows, threads=th) == -1
    n = rp.size - 1
    # grid closed form: row + col levels pass; the verifier agrees with them
    g_rp, g_col = oracle.gen_grid(16, 16)[1:]
    g_lv = np.array([(i // 16) + (i % 16) for i in range(256)], np.int32)
    assert oracle.bfs_verify(g_rp, g_col, np.array([0], np.uint32), g_lv[None]) == -1
    reached = np.flatnonzero((rows[2] != INF) & (rows[2] > 1))
    leaf = reached[0]
    plants = {
        "one too high": lambda r: r.__setitem__(leaf, r[leaf] + 1),
        "one too low": lambda r: r.__setitem__(leaf, r[leaf] - 1),
        "reached set to INF": lambda r: r.__setitem__(leaf, INF),
        "second level 0": lambda r: r.__setitem__(leaf, 0),
        "source not 0": lambda r: r.__setitem__(int(srcs[2]), 1),
    }
    for name, plant in plants.items():
        bad = rows.copy()
        plant(bad[2])
        assert oracle.bfs_verify(rp, col, srcs, bad, threads=2) == 2, name
    # an unreachable vertex given a finite level
    unreached = np.flatnonzero(rows[3] == INF)
    if unreached.size:
        bad = rows.copy()
        bad[3][unreached[0]] = 3
        assert oracle.bfs_verify(rp, col, srcs, bad) == 3
    assert n == 2048


# -------------------------------------------------------------------- SSSP
def test_sssp_matches_networkx_golden():
    keys = [k for k in G.files if "/sssp/" in k]
    assert keys
    for k in keys:
        name, _, s = k.split("/")
        if name == "unif2k":
            rp, col, w = G["unif2k/row_ptr"], G["unif2k/col"], G["unif2k/w"]
        else:
            _, rp, col, w = load(name)
        d, rounds = oracle.sssp(rp, col, w, int(s))
        assert rounds >= 1
        assert np.array_equal(d, G[k]), k
        assert np.array_equal(oracle.dijkstra(rp, col, w, int(s)), G[k]), k


def test_sssp_unit_weights_equal_bfs():
    _, rp, col = oracle.gen_rmat(11, 8, seed=2)
    w = np.ones(col.size, np.int32)
    for s in (0, 5):
        assert np.array_equal(oracle.sssp(rp, col, w, s)[0], oracle.bfs(rp, col, s))


def test_sssp_frontier_equals_dijkstra_random():
    _, rp, col = oracle.gen_uniform(5000, 40000, seed=9)
    w = oracle.weights_i32(col.size, 1, 64, 9)
    for s in (0, 17, 4999):
        d1, _ = oracle.sssp(rp, col, w, s)
        d2 = oracle.dijkstra(rp, col, w, s)
        assert np.array_equal(d1, d2)
        n = rp.size - 1
        src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        fin = d1[src] != INF
        assert np.all(d1[col[fin]].astype(np.int64) <= d1[src[fin]].astype(np.int64) + w[fin])


def test_sssp_negative_cycle_detected():
    rp = np.array([0, 1, 2], np.uint32)
    col = np.array([1, 0], np.uint32)
    w = np.array([-1, -1], np.int32)
    _, rounds = oracle.sssp(rp, col, w, 0)
    assert rounds == -1


# ---------------------------------------------------------------- PageRank
def test_pagerank_star_closed_form():
    n, rp, col, _ = load("starGraph_1K")
    r, updates, _ = oracle.pagerank(rp, col, 20)
    # centre c and leaf l: c' = d*999*l + (1-d)/N, l' = d*c/999 + (1-d)/N
    c = l = 1.0 / n
    for _ in range(updates):
        c, l = 0.85 * 999 * l + 0.15 / n, 0.85 * c / 999 + 0.15 / n
    assert abs(r[0] - c) <= 1e-15 * c * 50 and np.allclose(r[1:], l, rtol=1e-13, atol=0)


def test_pagerank_isolated():
    n, rp, col, _ = load("isolatedGraph_1K")
    r, updates, _ = oracle.pagerank(rp, col, 20)
    assert np.allclose(r, 0.15 / n, rtol=1e-15)
    assert updates == 2  # second update has zero delta -> stops


def test_pagerank_matches_scipy_golden():
    keys = [k for k in G.files if k.endswith("/pr20") or k.endswith("/pr1000")]
    for k in keys:
        name, what = k.split("/")
        iters = int(what[2:])
        if name.startswith("rmat"):
            rp, col = G[f"{name}/row_ptr"], G[f"{name}/col"]
        else:
            _, rp, col, _ = load(name)
        r, _, _ = oracle.pagerank(rp, col, iters)
        np.testing.assert_allclose(r, G[k], rtol=1e-12, atol=0, err_msg=k)


def test_pagerank_mass_identity():
    _, rp, col = oracle.gen_rmat(10, 16, seed=7)
    n = rp.size - 1
    r1, _, _ = oracle.pagerank(rp, col, 1)
    r0 = np.full(n, 1.0 / n)
    deg = np.diff(rp.astype(np.int64))
    assert abs(r1.sum() - (0.85 * r0[deg > 0].sum() + 0.15)) < 1e-12


# ------------------------------------------------- weighted PR and adjoint
def small_weighted(seed=4, scale=7):
    n, rp, col = oracle.gen_rmat(scale, 8, seed=seed)
    w = oracle.weights_f32_normalized(rp, seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    t = rng.exponential(size=n)
    t /= t.sum()
    return n, rp, col, w, t


def dense_P(n, rp, col, w):
    P = np.zeros((n, n))
    src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    np.add.at(P, (src, col.astype(np.int64)), w)
    return P


def test_weighted_pagerank_dense_solve():
    n, rp, col, w, t = small_weighted()
    r, it = oracle.pagerank_weighted(rp, col, w, t, tol=1e-15)
    P = dense_P(n, rp, col, w)
    r_ref = np.linalg.solve(np.eye(n) - 0.85 * P.T, 0.15 * t)
    np.testing.assert_allclose(r, r_ref, rtol=1e-11, atol=1e-15)


def test_weighted_pagerank_reduces_to_reference_update():
    _, rp, col = oracle.gen_rmat(9, 8, seed=3)
    n = rp.size - 1
    deg = np.diff(rp.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    w = 1.0 / deg[src]
    t = np.full(n, 1.0 / n)
    r_ref, _, _ = oracle.pagerank(rp, col, 20, epsilon=0.0)
    # tol=0 never converges: exactly 20 updates through the iteration cap
    r, it = oracle.pagerank_weighted(rp, col, w, t, tol=0.0, max_iters=20, allow_cap=True)
    assert it == 20
    np.testing.assert_allclose(r, r_ref, rtol=1e-13)
    with pytest.raises(ValueError):
        oracle.pagerank_weighted(rp, col, w, t, tol=0.0, max_iters=20)


def test_adjoint_dense_and_fd():
    n, rp, col, w, t = small_weighted(seed=11, scale=6)
    rng = np.random.default_rng(0)
    gbar = rng.standard_normal(n)
    r, _ = oracle.pagerank_weighted(rp, col, w, t, tol=1e-15)
    gw, gt, lam = oracle.pagerank_vjp(rp, col, w, r, gbar)
    P = dense_P(n, rp, col, w)
    lam_ref = np.linalg.solve(np.eye(n) - 0.85 * P, gbar)
    np.testing.assert_allclose(lam, lam_ref, rtol=1e-10, atol=1e-12)
    # central finite differences of L = <gbar, r(w, t)> along random directions
    L = lambda ww, tt: float(gbar @ oracle.pagerank_weighted(rp, col, ww, tt, tol=1e-15)[0])
    for k in range(3):
        dw = rng.standard_normal(w.size)
        dt = rng.standard_normal(n)
        h = 1e-6
        fd = (L(w + h * dw, t + h * dt) - L(w - h * dw, t - h * dt)) / (2 * h)
        an = gw @ dw + gt @ dt
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_adjoint_plants_mistake():
    # dropping the damping factor from grad_w must fail the FD check
    n, rp, col, w, t = small_weighted(seed=12, scale=6)
    rng = np.random.default_rng(1)
    gbar = rng.standard_normal(n)
    r, _ = oracle.pagerank_weighted(rp, col, w, t, tol=1e-15)
    gw, gt, lam = oracle.pagerank_vjp(rp, col, w, r, gbar)
    bad = gw / 0.85
    dw = rng.standard_normal(w.size)
    h = 1e-6
    L = lambda ww: float(gbar @ oracle.pagerank_weighted(rp, col, ww, t, tol=1e-15)[0])
    fd = (L(w + h * dw) - L(w - h * dw)) / (2 * h)
    assert abs(fd - gw @ dw) < 1e-6 * max(1, abs(fd))
    assert abs(fd - bad @ dw) > 1e-4 * max(1, abs(fd))


def test_c_edge_list_loader_matches_python_restatement():
    """oracle.c's edge-list loader (the CPU baseline of the ingest workload)
    equals the Python restatement on the reference's data files and on
    generated text with comments, blank lines, CRLF, tabs and weights."""
    import olive_amd
    from conftest import DATA_FILES, data_file

    def same(a, b):
        return (a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
                and ((a[3] is None and b[3] is None) or np.array_equal(a[3], b[3])))

    for name in DATA_FILES:
        with open(data_file(name), "rb") as fh:
            text = fh.read()
        assert same(oracle.load_edge_list(data_file(name)), oracle.load_edge_list_text(text)), name
    rng = np.random.default_rng(2)
    n, m = 3000, 40000
    for w in (None, rng.integers(-10 ** 9, 10 ** 9, m)):
        text = olive_amd.format_edge_list(n, rng.integers(0, n, m), rng.integers(0, n, m), w)
        lines = text.split(b"\n")
        messy = b"# c\n\n" + b"\n".join((b"  " + ln.replace(b" ", b"\t") + b"\r") if i % 3 == 1 else ln
                                        for i, ln in enumerate(lines))
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".txt", delete=False) as fh:
            fh.write(messy)
            path = fh.name
        try:
            assert same(oracle.load_edge_list(path), oracle.load_edge_list_text(messy))
        finally:
            os.unlink(path)


def test_print_values_is_printf(tmp_path):
    """oracle.print_values (the egress CPU baseline) writes what Python's
    "%d" and libc's "%f" formatting give."""
    import ctypes
    libc = ctypes.CDLL(None)
    libc.snprintf.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_double]
    a = np.array([0, -1, 7, 2 ** 31 - 1, -2 ** 31], np.int32)
    p = tmp_path / "i.txt"
    oracle.print_values(p, a)
    assert p.read_bytes() == "".join(f"{x}\n" for x in a.tolist()).encode()
    x = np.array([0.0, -0.0, 1e-7, 0.0078125, 123.4567895, -2.5e-7])
    oracle.print_values(p, x)
    buf = ctypes.create_string_buffer(64)
    want = b""
    for v in x:
        n = libc.snprintf(buf, 64, b"%f\n", float(v))
        want += buf.raw[:n]
    assert p.read_bytes() == want
