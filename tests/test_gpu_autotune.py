"""The autotuner (olv_autotune_ex): a tuned graph keeps every result bit
for bit (BFS in both numberings, batched BFS and its plan, SSSP, PageRank
row pull); the result is stored under the graph's signature and a second
handle of the same graph, or a plan of a never-tuned handle with
OLV_TUNE_CACHE set, takes it from the cache; invalid settings are refused."""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_autotune_exact_and_cached(olv, tmp_path):
    from olive_amd import _lib
    L = _lib.lib()
    csr = olv.with_weights_i32(olv.gen_rmat(14, 16, seed=2), 1, 64, seed=2)
    g = olv.Graph(csr)
    cache = tmp_path / "tune.jsonl"
    t = g.autotune_ex(_lib.OLV_TUNE_ALL, cache)
    say("tuned:", t.as_dict())
    assert t.source == 1 and t.tune_ms > 0
    assert g.signature() == olv.csr_signature(csr)
    nz = np.flatnonzero(csr.outdeg())
    srcs = nz[[0, 7, 100, 2000, len(nz) - 1]].astype(np.uint32)
    for s in srcs:
        ref = oracle.bfs(csr.row_ptr, csr.col, int(s))
        for fl in (0, _lib.OLV_GRAPH_ORDER, _lib.OLV_NO_DIRECTION_OPT):
            assert np.array_equal(g.bfs(int(s), flags=fl)[0].numpy(), ref), (s, fl)
    out, _ = g.bfs_batched(srcs)
    got = out.numpy().reshape(len(srcs), csr.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(s)))
    plan = ctypes.c_void_p()
    pout = olv.DeviceBuffer(csr.n * len(srcs), np.int32)
    _lib.check(L.olv_plan_bfs_batched(g.h, srcs.ctypes.data, len(srcs), pout.c, 0, ctypes.byref(plan)))
    _lib.check(L.olv_plan_launch(plan, None))
    assert np.array_equal(pout.numpy(), out.numpy())
    L.olv_plan_destroy(plan)
    assert np.array_equal(g.sssp(int(srcs[1]))[0].numpy(),
                          oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, int(srcs[1])))
    a = g.pagerank(20, epsilon=0.0, flags=_lib.OLV_PR_ROW_PULL)[0].numpy()
    b = g.pagerank(20, epsilon=0.0, flags=_lib.OLV_GRAPH_ORDER)[0].numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, 20, epsilon=0.0)
    assert float(np.max(np.abs(a - ref) / ref)) <= 1e-5
    # a second handle takes the stored result without measuring
    g2 = olv.Graph(csr)
    t2 = g2.autotune_ex(_lib.OLV_TUNE_ALL, cache)
    assert t2.source == 2 and t2.tune_ms == 0
    d1, d2 = t.as_dict(), t2.as_dict()
    for k in d1:
        if k not in ("source", "tune_ms"):
            assert d1[k] == pytest.approx(d2[k]), k
    # a plan of a never-tuned handle reads the cache named by OLV_TUNE_CACHE
    g3 = olv.Graph(csr)
    assert g3.tuning().source == 0
    os.environ["OLV_TUNE_CACHE"] = str(cache)
    try:
        one = np.array([srcs[2]], np.uint32)
        lv = olv.DeviceBuffer(csr.n, np.int32)
        _lib.check(L.olv_plan_bfs(g3.h, one.ctypes.data, 1, lv.c, 0, ctypes.byref(plan)))
        _lib.check(L.olv_plan_launch(plan, None))
        L.olv_plan_destroy(plan)
        assert g3.tuning().source == 2
        assert np.array_equal(lv.numpy(), oracle.bfs(csr.row_ptr, csr.col, int(srcs[2])))
    finally:
        del os.environ["OLV_TUNE_CACHE"]


def test_set_tuning_validates(olv):
    from olive_amd import _lib
    csr = olv.gen_rmat(10, 16, seed=3)
    g = olv.Graph(csr)
    t = g.tuning()
    t.seg_unroll = 16
    t.hub_thr = 512
    t.seg_bounds[:] = [4, 48, 512]
    g.set_tuning(t)
    got = g.tuning()
    assert got.source == 3 and got.seg_unroll == 16 and list(got.seg_bounds) == [4, 48, 512]
    s = int(np.argmax(csr.outdeg()))
    assert np.array_equal(g.bfs(s)[0].numpy(), oracle.bfs(csr.row_ptr, csr.col, s))
    ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, 20, epsilon=0.0)
    a = g.pagerank(20, epsilon=0.0, flags=_lib.OLV_PR_ROW_PULL)[0].numpy()
    assert float(np.max(np.abs(a - ref) / ref)) <= 1e-5
    for field, bad in (("seg_unroll", 3), ("hub_thr", 128), ("td_wave_thr", 300)):
        t = g.tuning()
        setattr(t, field, bad)
        with pytest.raises(olv.OliveError) as e:
            g.set_tuning(t)
        assert e.value.code == 1 and field.split("_")[0] in str(e.value)


def test_sssp_step_changes_rounds_not_distances(olv):
    """The near-far step (sssp_delta_num) shapes the rounds of both SSSP
    paths; every setting gives Dijkstra's distances, and an out-of-range one
    is refused."""
    from olive_amd import _lib
    csr = olv.with_weights_i32(olv.gen_uniform(1 << 15, 1 << 18, seed=8), 1, 64, seed=8)
    g = olv.Graph(csr, transpose=False)
    ref = oracle.dijkstra(csr.row_ptr, csr.col, csr.weights, 5)
    rounds = {}
    for num in (1, 8, 32, 256, 1024):
        t = g.tuning()
        t.sssp_delta_num = num
        g.set_tuning(t)
        for fl in (0, _lib.OLV_SSSP_DIRECT):
            out, st = g.sssp(5, flags=fl)
            assert np.array_equal(out.numpy(), ref), (num, fl)
            rounds[(num, fl)] = st.iterations
    say("rounds by step:", rounds)
    assert rounds[(1, 0)] > rounds[(1024, 0)]
    t = g.tuning()
    t.sssp_delta_num = 0
    with pytest.raises(olv.OliveError) as e:
        g.set_tuning(t)
    assert e.value.code == 1 and "sssp_delta_num" in str(e.value)
