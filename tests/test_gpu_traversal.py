"""GPU parity for BFS and SSSP through the C ABI: bit-exact against the
oracle (and the networkx golden vectors) on the reference's data files,
the grid, RMAT and uniform graphs, in every driver mode."""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import DATA_FILES, GOLDEN, data_file

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF
G = np.load(os.path.join(GOLDEN, "golden.npz"))

MODES = {
    "pipeline": 0,                      # device-controlled level kernels
    "host_loop": 0x8,                   # host waits for every level
    "push_only": 0x2,                   # no bottom-up steps
    "push_only_host_loop": 0x2 | 0x8,
    "persistent": 0x20,                 # one cooperative kernel, grid barrier
    "persistent_push_only": 0x20 | 0x2,
    "bottom_up": 0x4,                   # every level after the first pulls
    "persistent_bottom_up": 0x20 | 0x4,
    "direct": 0x800,                    # SSSP: per-edge probe + atomics, no binning
    "direct_host_loop": 0x800 | 0x8,
    # the same drivers in the graph's own numbering (the others run on the
    # graph renumbered by descending in-degree)
    "graph_order": 0x40,
    "graph_order_host_loop": 0x40 | 0x8,
    "graph_order_persistent": 0x40 | 0x20,
    "graph_order_bottom_up": 0x40 | 0x4,
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", DATA_FILES)
def test_bfs_data_files(olv, name, mode):
    say(f"bfs {name} {mode}")
    csr = olv.load_edge_list(data_file(name))
    g = olv.Graph(csr)
    srcs = [0] if csr.n < 3 else [0, csr.n // 2, csr.n - 1]
    for s in srcs:
        out, st = g.bfs(s, flags=MODES[mode])
        got = out.numpy()
        assert np.array_equal(got, G[f"{name}/bfs/{s}"]), (name, s)
        assert np.array_equal(got, oracle.bfs(csr.row_ptr, csr.col, s))
        # one expansion per level plus the final empty one
        assert st.iterations == got[got != INF].max() + 1


@pytest.mark.parametrize("mode", list(MODES))
def test_bfs_grid128(olv, mode):
    say(f"bfs grid128 {mode}")
    csr = olv.gen_grid(128, 128)
    g = olv.Graph(csr)
    out, st = g.bfs(0, flags=MODES[mode])
    i = np.arange(csr.n)
    assert np.array_equal(out.numpy(), (i // 128 + i % 128).astype(np.int32))
    assert st.iterations == 255  # 254 levels + the empty one


@pytest.mark.parametrize("scale,seed", [(12, 1), (16, 2), (18, 3)])
def test_bfs_rmat(olv, scale, seed):
    say(f"bfs rmat{scale}")
    csr = olv.gen_rmat(scale, 16, seed=seed)
    g = olv.Graph(csr)
    deg = csr.outdeg()
    srcs = [int(np.argmax(deg)), 0, int(np.flatnonzero(deg)[len(np.flatnonzero(deg)) // 2])]
    for s in srcs:
        ref = oracle.bfs(csr.row_ptr, csr.col, s)
        for mode, fl in MODES.items():
            out, st = g.bfs(s, flags=fl)
            assert np.array_equal(out.numpy(), ref), (scale, s, mode)
        out, st = g.bfs(s)
        reached = ref != INF
        assert st.vertices_reached == reached.sum()
        assert st.edges_traversed == deg[reached].sum()


def test_bfs_rmat22_full_size(olv):
    say("bfs rmat22")
    csr = olv.gen_rmat(22, 16, seed=1)
    g = olv.Graph(csr)
    s = int(np.argmax(csr.outdeg()))
    ref = oracle.bfs(csr.row_ptr, csr.col, s)
    for fl in (0x8, 0x20, 0):
        out, st = g.bfs(s, flags=fl)
        assert np.array_equal(out.numpy(), ref), fl
    assert st.bottom_up_levels > 0


def test_bfs_deterministic_and_reusable(olv):
    csr = olv.gen_rmat(14, 16, seed=9)
    g = olv.Graph(csr)
    a = g.bfs(5)[0].numpy()
    for _ in range(5):
        assert np.array_equal(g.bfs(5)[0].numpy(), a)


def test_bfs_multisource(olv):
    csr = olv.gen_rmat(13, 8, seed=4)
    g = olv.Graph(csr)
    srcs = [3, 100, 5000, 3]
    ref = np.min(np.stack([oracle.bfs(csr.row_ptr, csr.col, s).astype(np.int64) for s in srcs]), axis=0)
    for fl in (0, 0x40):
        out, st = g.bfs_multisource(srcs, flags=fl)
        assert np.array_equal(out.numpy(), ref.astype(np.int32)), fl


def test_bfs_twin_matches_graph_order_stats(olv):
    """The renumbered twin gives the same levels and the same statistics as
    the run in the graph's own numbering, also on a graph with isolated
    vertices and a weighted one."""
    csr = olv.with_weights_i32(olv.gen_rmat(15, 4, seed=11), seed=11)
    g = olv.Graph(csr)
    for s in (0, 77, int(np.argmax(csr.outdeg()))):
        a, sa = g.bfs(s)
        b, sb = g.bfs(s, flags=0x40)
        assert np.array_equal(a.numpy(), b.numpy()), s
        assert (sa.iterations, sa.vertices_reached, sa.edges_traversed) == \
               (sb.iterations, sb.vertices_reached, sb.edges_traversed), s


def test_bfs_errors(olv):
    g = olv.Graph(olv.gen_grid(4, 4))
    with pytest.raises(olv.OliveError) as e:
        g.bfs(16)
    assert e.value.code == 1


# -------------------------------------------------------------------- SSSP
@pytest.mark.parametrize("mode", ["pipeline", "host_loop", "persistent", "direct"])
def test_sssp_weighted_data_files(olv, mode):
    for name in DATA_FILES:
        csr = olv.load_edge_list(data_file(name))
        if csr.weights is None:
            continue
        say(f"sssp {name} {mode}")
        g = olv.Graph(csr)
        for s in [0, csr.n // 2, csr.n - 1]:
            out, st = g.sssp(s, flags=MODES[mode])
            assert np.array_equal(out.numpy(), G[f"{name}/sssp/{s}"]), (name, s)


@pytest.mark.parametrize("mode", ["pipeline", "host_loop", "persistent", "direct", "direct_host_loop"])
def test_sssp_uniform(olv, mode):
    say(f"sssp uniform {mode}")
    csr = olv.with_weights_i32(olv.gen_uniform(1 << 16, 1 << 19, seed=7), 1, 64, seed=7)
    g = olv.Graph(csr)
    for s in (0, 777):
        ref = oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, s)
        out, st = g.sssp(s, flags=MODES[mode])
        assert np.array_equal(out.numpy(), ref)
        assert np.array_equal(oracle.sssp(csr.row_ptr, csr.col, csr.weights, s)[0], ref)


def test_sssp_golden_unif2k(olv):
    rp, col, w = G["unif2k/row_ptr"], G["unif2k/col"], G["unif2k/w"]
    g = olv.Graph(olv.CsrGraph(rp, col, w))
    for s in (0, 999):
        assert np.array_equal(g.sssp(s)[0].numpy(), G[f"unif2k/sssp/{s}"])


def test_sssp_unit_weights_equal_bfs(olv):
    csr = olv.gen_rmat(14, 16, seed=2)
    csr1 = csr.with_weights(np.ones(csr.m, np.int32))
    g = olv.Graph(csr1)
    for s in (0, int(np.argmax(csr.outdeg()))):
        assert np.array_equal(g.sssp(s)[0].numpy(), g.bfs(s)[0].numpy())


def test_sssp_rmat_heavy_weights(olv):
    csr = olv.with_weights_i32(olv.gen_rmat(15, 16, seed=5), 1, 1000, seed=5)
    g = olv.Graph(csr)
    s = int(np.argmax(csr.outdeg()))
    assert np.array_equal(g.sssp(s)[0].numpy(), oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, s))


def test_sssp_negative_weights_without_cycle(olv):
    """Negative weights (an acyclic graph) switch off the near-far split;
    distances still equal the oracle's frontier relaxation."""
    csr0 = olv.gen_rmat(12, 8, seed=9)
    src = np.repeat(np.arange(csr0.n), csr0.outdeg())
    keep = src < csr0.col  # forward edges only: a DAG
    rp = np.zeros(csr0.n + 1, np.uint32)
    np.cumsum(np.bincount(src[keep], minlength=csr0.n), out=rp[1:])
    col = csr0.col[keep].astype(np.uint32)
    w = (np.random.default_rng(3).integers(-20, 60, col.size)).astype(np.int32)
    csr = olv.CsrGraph(rp, col, w)
    g = olv.Graph(csr)
    for s in (0, int(np.argmax(csr.outdeg()))):
        ref, rounds = oracle.sssp(rp, col, w, s)
        assert rounds != -1
        assert np.array_equal(g.sssp(s)[0].numpy(), ref)


def test_sssp_zero_and_large_weights(olv):
    """Near-far split with zero weights (step 1) and with weights up to 2^20."""
    base = olv.gen_uniform(1 << 14, 1 << 17, seed=4)
    for lo, hi in ((0, 1), (0, 3), (1, 1 << 20)):
        csr = olv.with_weights_i32(base, lo, hi, seed=lo + hi)
        g = olv.Graph(csr, transpose=False)
        ref = oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, 0)
        assert np.array_equal(g.sssp(0)[0].numpy(), ref), (lo, hi)


@pytest.mark.parametrize("flags", [0, 0x8, 0x800, 0x808])
def test_sssp_chain_with_one_heavy_edge(olv, flags):
    """A 1000-vertex chain with unit weights and one weight-1e6 edge: with
    direct relaxation (0x800) the near-far step (~32k) needs ~31 promotions
    with nothing to relax, so the run takes more rounds than the chain has
    vertices; the binned path re-splits past the heavy edge at once."""
    n = 1000
    rp = np.arange(n + 1, dtype=np.uint32)
    rp[n] = n - 1
    col = np.arange(1, n, dtype=np.uint32)
    w = np.ones(n - 1, np.int32)
    w[500] = 1_000_000
    csr = olv.CsrGraph(rp, col, w)
    g = olv.Graph(csr, transpose=False)
    out, st = g.sssp(0, flags=flags)
    ref = oracle.dijkstra(rp, col, w, 0)
    assert ref[-1] == n - 2 + 1_000_000
    assert np.array_equal(out.numpy(), ref)
    say(f"  chain: {st.iterations} rounds")
    if not flags & 0x800:
        assert st.iterations <= n + 2  # one round per hop, plus one re-split
    assert st.edges_inspected == n - 1  # every edge relaxed exactly once


def test_sssp_negative_cycle_reports_enoconv(olv):
    csr = olv.CsrGraph(np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.array([-1, -1], np.int32))
    g = olv.Graph(csr)
    for fl in (0, 0x8, 0x20):
        with pytest.raises(olv.OliveError) as e:
            g.sssp(0, flags=fl)
        assert e.value.code == 6


def test_sssp_full_size_uniform(olv):
    say("sssp uniform 2^22 x 2^25")
    csr = olv.with_weights_i32(olv.gen_uniform(1 << 22, 1 << 25, seed=1), 1, 64, seed=1)
    g = olv.Graph(csr, transpose=False)
    ref = oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, 0)
    for fl in (0, 0x800):
        out, st = g.sssp(0, flags=fl)
        assert np.array_equal(out.numpy(), ref), fl
        say(f"  flags {fl:#x}: {st.iterations} rounds, {st.edges_inspected} edges read, {st.elapsed_ms:.3f} ms")
    # a second call reuses the graph's bins and must reset every round's state
    out2, _ = g.sssp(0)
    assert np.array_equal(out2.numpy(), ref)


def sources(csr, k, seed):
    """k distinct-ish sources among vertices with out-edges."""
    nz = np.flatnonzero(csr.outdeg())
    rng = np.random.default_rng(seed)
    return rng.choice(nz, size=k, replace=k > nz.size).astype(np.uint32)


def test_sssp_binned_hubs_and_instances(olv):
    """Binned relaxation with hub items (RMAT, out-degree >= hub_thr) and k
    instances whose virtual ranges straddle partitions (n not a multiple of
    the 8192-vertex slice), against Dijkstra; zero-weight edges included."""
    say("sssp binned: rmat hubs, 6 instances")
    csr = olv.with_weights_i32(olv.gen_rmat(16, 16, seed=5), 0, 9, seed=5)
    assert int(csr.outdeg().max()) >= 1024
    g = olv.Graph(csr, transpose=False)
    srcs = sources(csr, 6, 11)
    out, st = g.sssp_batched(srcs)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i, s_ in enumerate(srcs):
        assert np.array_equal(got[i], oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, int(s_))), i
    one, _ = g.sssp(int(srcs[3]))
    assert np.array_equal(one.numpy(), got[3])
    # 260 instances: 17.0M virtual vertices, the expand/apply pair on
    # 32K-vertex slices
    say("sssp binned: rmat hubs, 260 instances")
    srcs = sources(csr, 260, 12)
    out, st = g.sssp_batched(srcs)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i in (0, 129, 259):
        assert np.array_equal(got[i], oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, int(srcs[i]))), i


def test_sssp_binned_wide_slices(olv):
    """k instances of a 300,007-vertex graph: 12.0M virtual vertices (one
    fused run), 24.0M and 60.0M (past the fused kernel's 16M: groups of 55
    instances, the last one shorter)."""
    csr = olv.with_weights_i32(olv.gen_uniform(300_007, 600_000, seed=9), 1, 100, seed=9)
    g = olv.Graph(csr, transpose=False)
    for k in (40, 80, 200):
        say(f"sssp binned wide slices: {k} instances")
        srcs = sources(csr, k, k)
        out, st = g.sssp_batched(srcs)
        got = out.numpy().reshape(k, csr.n)
        for i in (0, k // 2, k - 1):
            assert np.array_equal(got[i], oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, int(srcs[i]))), (k, i)
