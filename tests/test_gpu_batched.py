"""GPU parity for the batched paths: bit-parallel multi-source BFS, k-source
SSSP, the multi-graph union BFS and hipGraph plan replays -- each equal,
bit for bit, to per-source / per-graph oracle runs."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sources(csr, k, seed):
    """k sources among vertices with out-edges (Graph500 convention)."""
    nz = np.flatnonzero(csr.outdeg())
    rng = np.random.default_rng(seed)
    return rng.choice(nz, size=k, replace=k > nz.size).astype(np.uint32)


@pytest.mark.parametrize("k", [1, 37, 256, 300])
def test_bfs_batched_rmat(olv, k):
    say(f"batched bfs k={k}")
    csr = olv.gen_rmat(12, 16, seed=4)
    g = olv.Graph(csr)
    srcs = sources(csr, k, k)
    out, st = g.bfs_batched(srcs)
    got = out.numpy().reshape(k, csr.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(s))), (k, i, s)


@pytest.mark.parametrize("flags", [0, 0x2, 0x40])  # 0x40: the graph's own numbering
def test_bfs_batched_duplicates_and_grid(olv, flags):
    csr = olv.gen_grid(40, 50)
    g = olv.Graph(csr)
    srcs = np.array([0, 0, 1999, 17, 17, 1000], np.uint32)
    out, _ = g.bfs_batched(srcs, flags=flags)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(s)))


@pytest.mark.parametrize("flags", [0, 0x2, 0x40])
def test_bfs_batched_deep(olv, flags):
    """Levels of 255 and more (a 3 x 300 grid is 301 levels deep from a
    corner) next to shallow ones in the same rows."""
    say(f"batched bfs deep grid flags={flags}")
    csr = olv.gen_grid(3, 300)
    g = olv.Graph(csr)
    srcs = np.array([0, 899, 450, 299, 1, 0, 600], np.uint32)
    out, st = g.bfs_batched(srcs, flags=flags)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(s))), (i, s)
    assert got.max() >= 300


def test_bfs_batched_rmat18_256(olv):
    say("batched bfs rmat18 x 256")
    csr = olv.gen_rmat(18, 16, seed=1)
    g = olv.Graph(csr)
    srcs = sources(csr, 256, 1)
    out, st = g.bfs_batched(srcs)
    got = out.numpy().reshape(256, csr.n)
    for i in range(0, 256, 17):
        assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(srcs[i]))), i
    # on the renumbered twin (default) and in the graph's own numbering
    own, _ = g.bfs_batched(srcs, flags=0x40)
    assert np.array_equal(own.numpy().reshape(256, csr.n), got)


def test_sssp_batched(olv):
    say("batched sssp")
    csr = olv.with_weights_i32(olv.gen_uniform(20000, 160000, seed=2), 1, 64, seed=2)
    g = olv.Graph(csr, transpose=False)
    srcs = sources(csr, 24, 3)
    out, st = g.sssp_batched(srcs)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(got[i], oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, int(s)))
    for fl in (0x8, 0x800, 0x808):
        out2, _ = g.sssp_batched(srcs, flags=fl)
        assert np.array_equal(out2.numpy(), out.numpy()), fl


def test_multigraph_union(olv):
    say("multigraph")
    import ctypes
    from olive_amd import _lib
    csrs = [olv.gen_rmat(olv.batch_scale(5, i, 8, 12), 16, seed=olv.batch_seed(5, i)) for i in range(9)]
    gs = [olv.Graph(c) for c in csrs]
    srcs = np.array([int(np.flatnonzero(c.outdeg())[0]) for c in csrs], np.uint32)
    outs = [olv.DeviceBuffer(c.n, np.int32) for c in csrs]
    arr_g = (ctypes.c_void_p * len(gs))(*[g.h.value for g in gs])
    arr_o = (ctypes.c_void_p * len(outs))(*[o.ptr for o in outs])
    st = _lib.Stats()
    # direction-optimising, and bottom-up on every level after the first: the
    # union's in-CSR (and its in-list heads) is assembled on the device
    for flags in (0, _lib.OLV_FORCE_BOTTOM_UP):
        for o in outs:
            o.fill(0x5A5A5A5A)
        _lib.check(_lib.lib().olv_bfs_multigraph(arr_g, len(gs), srcs.ctypes.data, arr_o, flags, ctypes.byref(st)))
        for c, s, o in zip(csrs, srcs, outs):
            assert np.array_equal(o.numpy(), oracle.bfs(c.row_ptr, c.col, int(s))), flags


def test_plans_replay(olv):
    say("plans")
    import ctypes
    from olive_amd import _lib
    L = _lib.lib()
    csr = olv.gen_rmat(14, 16, seed=6)
    g = olv.Graph(csr)
    srcs = sources(csr, 64, 6)
    # plain multi-source plan
    out = olv.DeviceBuffer(csr.n, np.int32)
    plan = ctypes.c_void_p()
    one = np.array([srcs[0]], np.uint32)
    _lib.check(L.olv_plan_bfs(g.h, one.ctypes.data, 1, out.c, 0, ctypes.byref(plan)))
    ref = oracle.bfs(csr.row_ptr, csr.col, int(srcs[0]))
    for _ in range(3):
        out.fill(0)
        st = _lib.Stats()
        _lib.check(L.olv_plan_launch(plan, ctypes.byref(st)))
        assert np.array_equal(out.numpy(), ref)
        # eager work between replays
        assert np.array_equal(g.bfs(int(srcs[1]))[0].numpy(), oracle.bfs(csr.row_ptr, csr.col, int(srcs[1])))
    L.olv_plan_destroy(plan)
    # batched plan
    outb = olv.DeviceBuffer(csr.n * len(srcs), np.int32)
    planb = ctypes.c_void_p()
    _lib.check(L.olv_plan_bfs_batched(g.h, srcs.ctypes.data, len(srcs), outb.c, 0, ctypes.byref(planb)))
    eager = g.bfs_batched(srcs)[0].numpy()
    for _ in range(3):
        outb.fill(0)
        _lib.check(L.olv_plan_launch(planb, None))
        assert np.array_equal(outb.numpy(), eager)
    L.olv_plan_destroy(planb)


def test_autotune_keeps_results(olv):
    csr = olv.gen_rmat(14, 16, seed=2)
    g = olv.Graph(csr)
    s = int(np.argmax(csr.outdeg()))
    ref = oracle.bfs(csr.row_ptr, csr.col, s)
    tm, wm, bt = g.autotune()
    assert 1 <= tm < wm
    assert np.array_equal(g.bfs(s)[0].numpy(), ref)
