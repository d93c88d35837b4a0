"""GPU parity at the benchmark's full sizes.

* 256-source batched BFS on RMAT-22 (seed 1, sources drawn as the benchmark
  draws them) through a captured plan (olv_plan_bfs_batched): sampled rows
  bit-exact against the oracle's BFS (testBFS.cu:56-79), every row checked
  by the oracle's BFS-level verifier, every row equal to the run in the
  graph's own numbering, and rows equal to plain olv_bfs calls.
* The 64 RMAT graphs of the benchmark's batch (scales 12..18) as one graph
  union traversed by a captured plan (olv_plan_bfs): every graph bit-exact
  against the oracle, on two replays.
* RMAT-24 (16,777,216 vertices, past the 65535 x 256 = 16,776,960 threads
  of the reference's launch cap, utils.h:56-72 with oliverKernel.h:99-100,
  163-164): BFS bit-exact against the oracle, on the generator's ids and on
  a seeded scramble of the ids that puts reached vertices above
  16,776,960; PageRank (20 iterations) within 1e-5 relative of the fp64
  oracle at every vertex; the weighted forward, VJP and JVP checked in fp64
  on the host by full-size properties (fixed-point residuals, the gradient
  formulas, and the dot-product identity <gbar, J d> = <J^T gbar, d>).

Tolerances (stated, not fitted):
  * BFS levels: bit-exact
  * PageRank, 20 fixed iterations: max relative error <= 1e-5 vs fp64
  * weighted forward at tol 1e-6 (L1 change): fp64 residual
    ||r - d P^T r - (1-d) t||_1 <= 2e-6 (the stop rule bounds it by d * tol
    plus fp32 rounding)
  * adjoint at tol 1e-6 (relative to ||gbar||_1): fp64 residual
    ||lam - gbar - d P lam||_1 <= 4e-6 ||gbar||_1
  * gradients: relative L2 error of grad_w against d r_u lam_v and of grad_t
    against (1-d) lam <= 1e-5 (same r and lam: rounding only)
  * dot-product identity (JVP at tol 1e-7 on its L1 change, along a
    direction whose product does not cancel): relative difference <= 1e-4
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF
LAUNCH_CAP = 65535 * 256  # vertices the reference's kernelConfig reaches


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sources_of(csr, k, seed):
    """k sources among vertices with out-edges (the benchmark's draw)."""
    nz = np.flatnonzero(csr.outdeg())
    rng = np.random.default_rng(seed)
    return rng.choice(nz, size=k, replace=k > nz.size).astype(np.uint32)


# ---------------------------------------------------------------- RMAT-22
@pytest.fixture(scope="module")
def rmat22(olv):
    say("generating RMAT-22 seed 1")
    csr = olv.gen_rmat(22, 16, seed=1)
    return csr, olv.Graph(csr)


def test_batched_plan_rmat22_256(olv, rmat22):
    from olive_amd import _lib
    L = _lib.lib()
    csr, g = rmat22
    n = csr.n
    srcs = sources_of(csr, 256, 4)
    out = olv.DeviceBuffer(n * 256, np.int32)
    plan = ctypes.c_void_p()
    say("plan: 256 sources on RMAT-22")
    _lib.check(L.olv_plan_bfs_batched(g.h, srcs.ctypes.data, 256, out.c, 0, ctypes.byref(plan)))
    try:
        for rep in range(2):
            out.fill(0x5A5A5A5A)
            _lib.check(L.olv_plan_launch(plan, None))
            got = out.numpy().reshape(256, n)
            say(f"  replay {rep}: verifying 256 rows")
            assert oracle.bfs_verify(csr.row_ptr, csr.col, srcs, got) == -1
        for i in range(0, 256, 16):
            assert np.array_equal(got[i], oracle.bfs(csr.row_ptr, csr.col, int(srcs[i]))), i
    finally:
        L.olv_plan_destroy(plan)
    say("  graph's own numbering")
    own, _ = g.bfs_batched(srcs, flags=olv.OLV_GRAPH_ORDER)
    assert np.array_equal(own.numpy().reshape(256, n), got)
    del own
    one = olv.DeviceBuffer(n, np.int32)
    for i in (0, 85, 170, 255):
        g.bfs(int(srcs[i]), out=one)
        assert np.array_equal(one.numpy(), got[i]), i


# -------------------------------------------------------- 64-graph batch
def test_multigraph_64_captured_plan(olv):
    from olive_amd import _lib
    L = _lib.lib()
    say("generating the 64-graph batch (scales 12..18)")
    csrs = [olv.gen_rmat(olv.batch_scale(1, i, 12, 18), 16, seed=olv.batch_seed(1, i)) for i in range(64)]
    assert {olv.batch_scale(1, i, 12, 18) for i in range(64)} <= set(range(12, 19))
    gs = [olv.Graph(c) for c in csrs]
    bsrc = np.array([int(np.flatnonzero(c.outdeg())[0]) for c in csrs], np.uint32)
    arr = (ctypes.c_void_p * 64)(*[x.h.value for x in gs])
    uni = ctypes.c_void_p()
    offs = np.zeros(65, np.uint32)
    _lib.check(L.olv_graph_union(arr, 64, ctypes.byref(uni), offs.ctypes.data))
    assert np.array_equal(np.diff(offs.astype(np.int64)), [c.n for c in csrs])
    seeds = (offs[:64] + bsrc).astype(np.uint32)
    out = olv.DeviceBuffer(int(offs[64]), np.int32)
    plan = ctypes.c_void_p()
    _lib.check(L.olv_plan_bfs(uni, seeds.ctypes.data, 64, out.c, 0, ctypes.byref(plan)))
    refs = [oracle.bfs(c.row_ptr, c.col, int(s)) for c, s in zip(csrs, bsrc)]
    try:
        for rep in range(2):
            out.fill(0x5A5A5A5A)
            _lib.check(L.olv_plan_launch(plan, None))
            o = out.numpy()
            for i in range(64):
                assert np.array_equal(o[offs[i]:offs[i + 1]], refs[i]), (rep, i)
            # eager work between replays
            assert np.array_equal(gs[rep].bfs(int(bsrc[rep]))[0].numpy(), refs[rep])
    finally:
        L.olv_plan_destroy(plan)
        L.olv_graph_destroy(uni)


# ---------------------------------------------------------------- RMAT-24
@pytest.fixture(scope="module")
def rmat24(olv):
    say("generating weighted RMAT-24 seed 1")
    csr = olv.with_weights_f32_normalized(olv.gen_rmat(24, 16, seed=1), seed=1)
    assert csr.n == 1 << 24 and csr.n > LAUNCH_CAP
    t = olv.gen_dirichlet(csr.n, seed=1)
    g = olv.Graph(csr)
    return csr, t, g


def test_bfs_rmat24(olv, rmat24):
    csr, _, g = rmat24
    s = int(sources_of(csr, 1, 24)[0])
    say(f"bfs rmat24 from {s}")
    ref = oracle.bfs(csr.row_ptr, csr.col, s)
    for flags in (0, olv.OLV_GRAPH_ORDER):
        lv, st = g.bfs(s, flags=flags)
        assert np.array_equal(lv.numpy(), ref), flags


def test_bfs_rmat24_scrambled_ids(olv, rmat24):
    """The same graph with its vertex ids scrambled by a seeded permutation,
    so that reached vertices lie above the reference's launch cap."""
    csr, _, _ = rmat24
    n = csr.n
    perm = np.random.default_rng(24).permutation(n).astype(np.uint32)
    src = np.repeat(perm, csr.outdeg())
    dst = perm[csr.col]
    say("scrambled rmat24: device ingest")
    g = olv.ingest_coo(n, src, dst)
    del src, dst
    h = g.csr
    s = int(perm[int(sources_of(csr, 1, 24)[0])])
    ref = oracle.bfs(h.row_ptr, h.col, s)
    high = np.count_nonzero(ref[LAUNCH_CAP:] != INF)
    say(f"  reached above {LAUNCH_CAP}: {high}")
    assert high > 0
    for flags in (0, olv.OLV_GRAPH_ORDER):
        lv, _ = g.bfs(s, flags=flags)
        assert np.array_equal(lv.numpy(), ref), flags


def test_pagerank_rmat24(olv, rmat24):
    csr, _, g = rmat24
    say("pagerank rmat24, 20 iterations")
    out, st = g.pagerank(20)
    ref, updates, _ = oracle.pagerank(csr.row_ptr, csr.col, 20)
    got = out.numpy().astype(np.float64)
    rel = np.abs(got - ref) / ref
    say(f"  max rel err {rel.max():.3e}, above the cap {rel[LAUNCH_CAP:].max():.3e}")
    assert st.iterations == updates == 20
    assert rel.max() <= 1e-5


def _edge_sources(csr):
    return np.repeat(np.arange(csr.n, dtype=np.uint32), csr.outdeg())


def test_weighted_forward_vjp_jvp_rmat24(olv, rmat24):
    csr, t, g = rmat24
    n, m, d = csr.n, csr.m, 0.85
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    rng = np.random.default_rng(1)
    gbar = rng.standard_normal(n).astype(np.float32)
    say("weighted forward rmat24, tol 1e-6")
    r, st = g.pagerank_weighted(W, T, tol=1e-6)
    say(f"  {st.iterations} iterations, residual {st.residual:.3e}")
    src = _edge_sources(csr)
    col = csr.col
    w64 = csr.weights.astype(np.float64)
    r64 = r.numpy().astype(np.float64)
    t64 = t.astype(np.float64)
    y = np.bincount(col, weights=w64 * r64[src], minlength=n)
    res_f = float(np.abs(r64 - d * y - (1 - d) * t64).sum())
    say(f"  fp64 fixed-point residual {res_f:.3e}")
    assert res_f <= 2e-6
    # vertices with no in-edges hold exactly the teleport term
    noin = np.bincount(col, minlength=n) == 0
    assert np.allclose(r64[noin], (1 - d) * t64[noin], rtol=1e-6, atol=0)
    assert noin[LAUNCH_CAP:].any()
    del y
    say("vjp rmat24, tol 1e-6")
    gw, gt, lam, sa = g.pagerank_vjp(W, r, olv.DeviceBuffer.from_numpy(gbar), tol=1e-6)
    say(f"  {sa.iterations} iterations")
    lam64 = lam.numpy().astype(np.float64)
    g64 = gbar.astype(np.float64)
    z = np.bincount(src, weights=w64 * lam64[col], minlength=n)
    res_a = float(np.abs(lam64 - g64 - d * z).sum())
    say(f"  fp64 adjoint residual {res_a:.3e} (||gbar||_1 = {np.abs(g64).sum():.3e})")
    assert res_a <= 4e-6 * np.abs(g64).sum()
    del z
    gw64 = gw.numpy().astype(np.float64)
    gw_ref = d * r64[src] * lam64[col]
    rel_w = np.linalg.norm(gw64 - gw_ref) / np.linalg.norm(gw_ref)
    gt64 = gt.numpy().astype(np.float64)
    rel_t = np.linalg.norm(gt64 - (1 - d) * lam64) / np.linalg.norm((1 - d) * lam64)
    say(f"  grad_w rel {rel_w:.2e} grad_t rel {rel_t:.2e}")
    assert rel_w <= 1e-5 and rel_t <= 1e-5
    del gw_ref
    # direction: relative weight noise, and a teleport change along gbar so
    # that <gbar, J d> does not cancel to a sum of rounding errors
    say("jvp rmat24 and the dot-product identity")
    dw = (rng.standard_normal(m) * csr.weights).astype(np.float32)
    dt = (gbar / n).astype(np.float32)
    dr, sj = g.pagerank_jvp(W, r, olv.DeviceBuffer.from_numpy(dw), olv.DeviceBuffer.from_numpy(dt), tol=1e-7)
    say(f"  {sj.iterations} iterations")
    lhs = float(g64 @ dr.numpy().astype(np.float64))
    rhs = float(gw64 @ dw.astype(np.float64) + gt64 @ dt.astype(np.float64))
    say(f"  <gbar, J d> = {lhs:.8e}, <J^T gbar, d> = {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-4 * max(abs(lhs), abs(rhs))
