"""GPU parity for PageRank (fp32 with fp32 accumulation) against the fp64
oracle, and for the weighted fixed point and its VJP against fp64 dense
solves and central finite differences.

Tolerances (stated, not fitted):
  * unweighted PageRank, fixed iterations: max relative error <= 1e-5
  * weighted fixed point at tol 1e-7: L1 error <= 2e-6, max rel err <= 1e-4
  * VJP: relative L2 error of grad_w and grad_t <= 1e-4; directional
    derivatives within 1e-4 (relative) of fp64 central differences
"""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import DATA_FILES, GOLDEN, data_file

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(GOLDEN, "golden.npz"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def maxrel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.abs(b)))


@pytest.mark.parametrize("name", DATA_FILES)
def test_pagerank_data_files(olv, name):
    say(f"pr {name}")
    csr = olv.load_edge_list(data_file(name))
    g = olv.Graph(csr)
    out, st = g.pagerank(20)
    ref, updates, err = oracle.pagerank(csr.row_ptr, csr.col, 20)
    assert maxrel(out.numpy(), ref) <= 1e-5
    assert st.iterations == updates


@pytest.mark.parametrize("scale", [10, 14, 20])
def test_pagerank_rmat(olv, scale):
    say(f"pr rmat{scale}")
    csr = olv.gen_rmat(scale, 16, seed=1)
    g = olv.Graph(csr)
    out, st = g.pagerank(20)
    ref, updates, err = oracle.pagerank(csr.row_ptr, csr.col, 20)
    assert st.iterations == updates == 20
    e = maxrel(out.numpy(), ref)
    say(f"  max rel err {e:.3e}")
    assert e <= 1e-5
    assert abs(st.residual - err) <= 1e-4 * err


def test_pagerank_converges_like_reference(olv):
    # with a large iteration cap the epsilon rule decides, as in testPageRank
    csr = olv.load_edge_list(data_file("starGraph_1K"))
    g = olv.Graph(csr)
    out, st = g.pagerank(1000, epsilon=1e-5)
    ref, updates, _ = oracle.pagerank(csr.row_ptr, csr.col, 1000, epsilon=1e-5)
    assert st.iterations == updates
    assert maxrel(out.numpy(), ref) <= 1e-5


def test_pagerank_golden_rmat10(olv):
    g = olv.Graph(olv.CsrGraph(G["rmat10/row_ptr"], G["rmat10/col"]))
    out, _ = g.pagerank(20)
    assert maxrel(out.numpy(), G["rmat10/pr20"]) <= 1e-5


def test_pagerank_bit_reproducible(olv):
    g = olv.Graph(olv.gen_rmat(16, 16, seed=3))
    a = g.pagerank(20)[0].numpy()
    b = g.pagerank(20)[0].numpy()
    assert np.array_equal(a, b)


def star_in_hub(olv, n):
    """Every vertex but 0 and n-1 points at 0 (in-degree n-2: a chunked
    row), 0 points at the even vertices below n-1, and n-1 has no edges."""
    hub_out = np.arange(2, n - 1, 2)
    src = np.concatenate([np.arange(1, n - 1), np.zeros(hub_out.size, np.int64)])
    dst = np.concatenate([np.zeros(n - 2, np.int64), hub_out])
    o = np.argsort(src, kind="stable")
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(src, minlength=n))
    return olv.CsrGraph(rp, dst[o].astype(np.uint32))


@pytest.mark.parametrize("case", ["rmat16", "rmat18", "grid", "star"])
def test_pagerank_relabel_bit_identical(olv, case):
    """The per-row pull over the in-degree relabelled CSR (OLV_PR_ROW_PULL)
    gives the same fp32 ranks, bit for bit, as the iteration in the graph's
    own order.  The default source-segmented pull adds the same terms in
    another fixed order (per 16384-source segment, then across segments):
    like the others it is within 1e-5 relative of fp64, and run to run it is
    bit-reproducible.  rmat18 spans 16 segments; the star's hub row (20009
    in-edges from two segments) is cut into pieces of at most 512."""
    csr = {"rmat16": lambda: olv.gen_rmat(16, 16, seed=7), "rmat18": lambda: olv.gen_rmat(18, 16, seed=2),
           "grid": lambda: olv.gen_grid(64, 64), "star": lambda: star_in_hub(olv, 20011)}[case]()
    g = olv.Graph(csr)
    row = olv._lib.OLV_PR_ROW_PULL
    for iters in (1, 2, 20):
        a = g.pagerank(iters, epsilon=0.0, flags=row)[0].numpy()
        b = g.pagerank(iters, epsilon=0.0, flags=olv._lib.OLV_GRAPH_ORDER)[0].numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (case, iters)
        c = g.pagerank(iters, epsilon=0.0)[0].numpy()
        c2 = g.pagerank(iters, epsilon=0.0)[0].numpy()
        assert np.array_equal(c.view(np.uint32), c2.view(np.uint32)), (case, iters)
        ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, iters, epsilon=0.0)
        assert maxrel(a, ref) <= 1e-5 and maxrel(c, ref) <= 1e-5, (case, iters)
    say(f"  {case}: segmented vs row pull max rel diff {maxrel(c, a.astype(np.float64)):.2e}")


def test_pagerank_segmented_epsilon_stop(olv):
    """The stop rule (L1 change below epsilon) in the segmented pull stops
    at the same iteration as the oracle."""
    csr = olv.gen_rmat(17, 16, seed=11)
    g = olv.Graph(csr)
    for eps in (1e-3, 1e-5):
        out, st = g.pagerank(1000, epsilon=eps)
        ref, updates, _ = oracle.pagerank(csr.row_ptr, csr.col, 1000, epsilon=eps)
        assert st.iterations == updates, eps
        assert maxrel(out.numpy(), ref) <= 1e-5


# ------------------------------------------------ weighted + adjoint
def weighted_case(olv, scale, seed):
    csr = olv.with_weights_f32_normalized(olv.gen_rmat(scale, 16, seed=seed), seed=seed)
    t = olv.gen_dirichlet(csr.n, seed=seed)
    return csr, t


def bits(buf):
    return buf.numpy().view(np.uint32)


@pytest.mark.parametrize("case", ["rmat14", "star"])
def test_weighted_vjp_jvp_relabel_bit_identical(olv, case):
    """Forward, VJP (grad_w, grad_t, lambda) and JVP by the per-row pull
    over the renumbered CSRs (OLV_PR_ROW_PULL) equal the runs in the graph's
    own numbering bit for bit.  The default source-segmented pulls add the
    same terms in another fixed order: bit-reproducible run to run, and
    within 1e-5 (ranks, relative) and 1e-4 (relative L2 of the gradients
    and the tangent) of the row pull."""
    if case == "rmat14":
        csr, t = weighted_case(olv, 14, 21)
    else:
        base = star_in_hub(olv, 20011)
        csr = olv.with_weights_f32_normalized(base, seed=4)
        t = olv.gen_dirichlet(csr.n, seed=4)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights.astype(np.float32))
    T = olv.DeviceBuffer.from_numpy(t.astype(np.float32))
    rng = np.random.default_rng(5)
    gb = olv.DeviceBuffer.from_numpy(rng.standard_normal(csr.n).astype(np.float32))
    # small tangents (the tangent solve stops on an absolute L1 change, so
    # its fp32 rounding floor scales with the tangent's size)
    dW = olv.DeviceBuffer.from_numpy((1e-3 * rng.standard_normal(csr.m)).astype(np.float32))
    dT = olv.DeviceBuffer.from_numpy((1e-3 * rng.standard_normal(csr.n)).astype(np.float32))
    res = {}
    row = olv._lib.OLV_PR_ROW_PULL
    for key, fl in (("row", row), ("graph", olv.OLV_GRAPH_ORDER), ("seg", 0), ("seg2", 0)):
        r, st = g.pagerank_weighted(W, T, tol=1e-6, flags=fl)
        gw, gt, lam, sv = g.pagerank_vjp(W, r, gb, tol=1e-6, flags=fl)
        dr, sj = g.pagerank_jvp(W, r, dW, dT, tol=1e-6, flags=fl)
        res[key] = ([x.numpy().copy() for x in (r, gw, gt, lam, dr)], (st.iterations, sv.iterations, sj.iterations))
    names = ("r", "grad_w", "grad_t", "lambda", "jvp")
    for k1, k2 in (("row", "graph"), ("seg", "seg2")):
        (a, ia), (b, ib) = res[k1], res[k2]
        assert ia == ib, (case, k1, k2)
        for name, x, y in zip(names, a, b):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (case, k1, k2, name)
    (a, _), (b, _) = res["seg"], res["row"]
    assert maxrel(a[0], b[0].astype(np.float64)) <= 1e-5, case
    for name, x, y in zip(names[1:], a[1:], b[1:]):
        rel = np.linalg.norm(x.astype(np.float64) - y) / np.linalg.norm(y.astype(np.float64))
        assert rel <= 1e-4, (case, name, rel)


def test_weighted_pagerank_forward(olv):
    csr, t = weighted_case(olv, 12, 5)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    r, st = g.pagerank_weighted(W, T, tol=1e-7)
    ref, _ = oracle.pagerank_weighted(csr.row_ptr, csr.col, csr.weights.astype(np.float64),
                                      t.astype(np.float64), tol=1e-15)
    got = r.numpy().astype(np.float64)
    assert np.abs(got - ref).sum() <= 2e-6
    assert maxrel(got, ref) <= 1e-4
    # dense fp64 solve agrees with the oracle it was checked against
    say(f"  weighted iters {st.iterations} L1 {np.abs(got - ref).sum():.2e}")


def test_weighted_reduces_to_plain(olv):
    csr = olv.gen_rmat(12, 16, seed=8)
    deg = csr.outdeg()
    src = np.repeat(np.arange(csr.n), deg)
    w = (1.0 / deg[src]).astype(np.float32)
    t = np.full(csr.n, 1.0 / csr.n, np.float32)
    g = olv.Graph(csr)
    r, st = g.pagerank_weighted(olv.DeviceBuffer.from_numpy(w), olv.DeviceBuffer.from_numpy(t), tol=1e-9,
                                max_iters=500)
    ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, 1000, epsilon=1e-13)
    assert maxrel(r.numpy(), ref) <= 1e-4


def test_vjp_against_fp64(olv):
    csr, t = weighted_case(olv, 12, 6)
    rng = np.random.default_rng(0)
    gbar = rng.standard_normal(csr.n).astype(np.float32)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    r, _ = g.pagerank_weighted(W, T, tol=1e-8)
    gw, gt, lam, st = g.pagerank_vjp(W, r, olv.DeviceBuffer.from_numpy(gbar), tol=1e-7)
    w64, t64, g64 = csr.weights.astype(np.float64), t.astype(np.float64), gbar.astype(np.float64)
    r64, _ = oracle.pagerank_weighted(csr.row_ptr, csr.col, w64, t64, tol=1e-15)
    gw_ref, gt_ref, lam_ref = oracle.pagerank_vjp(csr.row_ptr, csr.col, w64, r64, g64)
    rel = lambda a, b: np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b)
    say(f"  vjp iters {st.iterations} rel gw {rel(gw.numpy(), gw_ref):.2e} gt {rel(gt.numpy(), gt_ref):.2e}")
    assert rel(gt.numpy(), gt_ref) <= 1e-4
    assert rel(gw.numpy(), gw_ref) <= 1e-4
    # fp64 central differences of L = <gbar, r(w, t)> along random directions
    L = lambda ww, tt: float(g64 @ oracle.pagerank_weighted(csr.row_ptr, csr.col, ww, tt, tol=1e-15)[0])
    gwn, gtn = gw.numpy().astype(np.float64), gt.numpy().astype(np.float64)
    for _ in range(3):
        dw = rng.standard_normal(csr.m)
        dt = rng.standard_normal(csr.n)
        h = 1e-5
        fd = (L(w64 + h * dw, t64 + h * dt) - L(w64 - h * dw, t64 - h * dt)) / (2 * h)
        an = gwn @ dw + gtn @ dt
        assert abs(fd - an) <= 1e-4 * max(abs(fd), 1e-3), (fd, an)


def test_vjp_dense_scale12(olv):
    """Scale-12 instance of the RMAT-24 generator against a dense fp64 solve."""
    csr, t = weighted_case(olv, 12, 24)
    n = csr.n
    rng = np.random.default_rng(2)
    gbar = rng.standard_normal(n).astype(np.float32)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    r, _ = g.pagerank_weighted(W, olv.DeviceBuffer.from_numpy(t), tol=1e-8)
    gw, gt, lam, _ = g.pagerank_vjp(W, r, olv.DeviceBuffer.from_numpy(gbar), tol=1e-7)
    P = np.zeros((n, n))
    src = np.repeat(np.arange(n), csr.outdeg())
    np.add.at(P, (src, csr.col.astype(np.int64)), csr.weights.astype(np.float64))
    r_ref = np.linalg.solve(np.eye(n) - 0.85 * P.T, 0.15 * t.astype(np.float64))
    lam_ref = np.linalg.solve(np.eye(n) - 0.85 * P, gbar.astype(np.float64))
    assert np.abs(r.numpy() - r_ref).sum() <= 2e-6
    assert np.linalg.norm(lam.numpy() - lam_ref) / np.linalg.norm(lam_ref) <= 1e-4
    gw_ref = 0.85 * r_ref[src] * lam_ref[csr.col.astype(np.int64)]
    assert np.linalg.norm(gw.numpy() - gw_ref) / np.linalg.norm(gw_ref) <= 1e-4
    assert np.linalg.norm(gt.numpy() - 0.15 * lam_ref) / np.linalg.norm(0.15 * lam_ref) <= 1e-4


def test_jvp_dense_and_dot_product(olv):
    """Forward mode: the tangent dr for (dw, dt) against the dense fp64
    tangent solve (I - d P^T) dr = d dP^T r + (1-d) dt, and the dot-product
    test <gbar, J (dw, dt)> = <J^T gbar, (dw, dt)> with the VJP."""
    csr, t = weighted_case(olv, 12, 24)
    n = csr.n
    rng = np.random.default_rng(5)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    r, _ = g.pagerank_weighted(W, T, tol=1e-8)
    gbar = rng.standard_normal(n).astype(np.float32)
    gw, gt, _, _ = g.pagerank_vjp(W, r, olv.DeviceBuffer.from_numpy(gbar), tol=1e-7)
    src = np.repeat(np.arange(n), csr.outdeg())
    dst = csr.col.astype(np.int64)
    P = np.zeros((n, n))
    np.add.at(P, (src, dst), csr.weights.astype(np.float64))
    r_ref = np.linalg.solve(np.eye(n) - 0.85 * P.T, 0.15 * t.astype(np.float64))
    for k in range(3):
        dw = rng.standard_normal(csr.m).astype(np.float32)
        dt = rng.standard_normal(n).astype(np.float32)
        dr, st = g.pagerank_jvp(W, r, olv.DeviceBuffer.from_numpy(dw), olv.DeviceBuffer.from_numpy(dt), tol=1e-7)
        dP = np.zeros((n, n))
        np.add.at(dP, (src, dst), dw.astype(np.float64))
        rhs = 0.85 * dP.T @ r_ref + 0.15 * dt.astype(np.float64)
        dr_ref = np.linalg.solve(np.eye(n) - 0.85 * P.T, rhs)
        got = dr.numpy().astype(np.float64)
        rel = np.linalg.norm(got - dr_ref) / np.linalg.norm(dr_ref)
        lhs = float(gbar.astype(np.float64) @ got)
        rhs_dot = float(gw.numpy().astype(np.float64) @ dw + gt.numpy().astype(np.float64) @ dt)
        say(f"  jvp {k}: iters {st.iterations} rel {rel:.2e} dot {lhs:.6e} vs {rhs_dot:.6e}")
        assert rel <= 1e-4
        assert abs(lhs - rhs_dot) <= 1e-4 * max(abs(lhs), abs(rhs_dot))


def test_jvp_against_central_differences(olv):
    """dr against fp64 central differences of the oracle's fixed point."""
    csr, t = weighted_case(olv, 12, 6)
    rng = np.random.default_rng(7)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    r, _ = g.pagerank_weighted(W, olv.DeviceBuffer.from_numpy(t), tol=1e-8)
    w64, t64 = csr.weights.astype(np.float64), t.astype(np.float64)
    dw = rng.standard_normal(csr.m).astype(np.float32)
    dt = rng.standard_normal(csr.n).astype(np.float32)
    dr, _ = g.pagerank_jvp(W, r, olv.DeviceBuffer.from_numpy(dw), olv.DeviceBuffer.from_numpy(dt), tol=1e-7)
    h = 1e-5
    f = lambda s: oracle.pagerank_weighted(csr.row_ptr, csr.col, w64 + s * dw, t64 + s * dt, tol=1e-15)[0]
    fd = (f(h) - f(-h)) / (2 * h)
    rel = np.linalg.norm(dr.numpy().astype(np.float64) - fd) / np.linalg.norm(fd)
    say(f"  jvp vs fd rel {rel:.2e}")
    assert rel <= 1e-4


def test_weighted_slot_weights_reuse(olv):
    """OLV_W_UNCHANGED reuses the weights already in slot order: results are
    bit-identical to a fresh permutation; weights rewritten in place and
    passed without the flag are re-permuted, and the other solver's copy of
    them is dropped, so a later flagged call does not see the old weights."""
    from olive_amd import _lib
    U = _lib.OLV_W_UNCHANGED
    csr, t = weighted_case(olv, 12, 24)
    g = olv.Graph(csr)
    W = olv.DeviceBuffer.from_numpy(csr.weights)
    T = olv.DeviceBuffer.from_numpy(t)
    G = olv.DeviceBuffer.from_numpy(np.random.default_rng(3).standard_normal(csr.n).astype(np.float32))
    D = olv.DeviceBuffer.from_numpy(np.random.default_rng(4).standard_normal(csr.m).astype(np.float32))
    r0, _ = g.pagerank_weighted(W, T, tol=1e-8)
    r1, _ = g.pagerank_weighted(W, T, tol=1e-8, flags=U)
    assert np.array_equal(bits(r0), bits(r1))
    a = g.pagerank_vjp(W, r0, G, tol=1e-7)
    b = g.pagerank_vjp(W, r0, G, tol=1e-7, flags=U)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(bits(x), bits(y))
    j0, _ = g.pagerank_jvp(W, r0, D, T, tol=1e-7)
    j1, _ = g.pagerank_jvp(W, r0, D, T, tol=1e-7, flags=U)
    assert np.array_equal(bits(j0), bits(j1))
    # new weights in the same buffer
    w2 = olv.with_weights_f32_normalized(olv.gen_rmat(12, 16, seed=24), seed=99).weights
    _lib.check(_lib.lib().olv_memcpy_h2d(W.ptr, w2.ctypes.data, W.nbytes))
    W2 = olv.DeviceBuffer.from_numpy(w2)
    r2, _ = g.pagerank_weighted(W, T, tol=1e-8)
    r2ref, _ = g.pagerank_weighted(W2, T, tol=1e-8)
    assert np.array_equal(bits(r2), bits(r2ref))
    assert not np.array_equal(bits(r2), bits(r0))
    c = g.pagerank_vjp(W, r2, G, tol=1e-7, flags=U)
    cref = g.pagerank_vjp(W2, r2, G, tol=1e-7)
    for x, y in zip(c[:3], cref[:3]):
        assert np.array_equal(bits(x), bits(y))
