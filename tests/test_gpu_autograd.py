"""The torch.autograd wrapper of weighted PageRank (olive_amd.autograd):
gradients from loss.backward() against a dense fp64 solve and fp64 central
finite differences of the oracle's fixed point, and the forward against the
oracle.  Tolerances: relative L2 error <= 1e-4 for gradients, forward L1
error <= 2e-6 (fp32 fixed point at tol 1e-8)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(olv, scale, seed):
    csr = olv.with_weights_f32_normalized(olv.gen_rmat(scale, 16, seed=seed), seed=seed)
    t = olv.gen_dirichlet(csr.n, seed=seed).astype(np.float32)
    return csr, t


def _dense(csr, t, gbar, d=0.85):
    n = csr.n
    src = np.repeat(np.arange(n), csr.outdeg())
    dst = csr.col.astype(np.int64)
    P = np.zeros((n, n))
    np.add.at(P, (src, dst), csr.weights.astype(np.float64))
    r = np.linalg.solve(np.eye(n) - d * P.T, (1 - d) * t.astype(np.float64))
    lam = np.linalg.solve(np.eye(n) - d * P, gbar.astype(np.float64))
    return r, d * r[src] * lam[dst], (1 - d) * lam


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_autograd_matches_dense_fp64(olv):
    from olive_amd import autograd

    csr, t = _case(olv, 11, 24)
    g = olv.Graph(csr)
    dev = torch.device("cuda:0")
    w = torch.tensor(csr.weights, device=dev, requires_grad=True)
    tt = torch.tensor(t, device=dev, requires_grad=True)
    gbar = np.random.default_rng(5).standard_normal(csr.n).astype(np.float32)
    r = autograd.pagerank(g, w, tt, tol=1e-8)
    loss = (r * torch.tensor(gbar, device=dev)).sum()
    loss.backward()
    r_ref, gw_ref, gt_ref = _dense(csr, t, gbar)
    assert np.abs(r.detach().cpu().numpy() - r_ref).sum() <= 2e-6
    assert _rel(w.grad.cpu().numpy(), gw_ref) <= 1e-4
    assert _rel(tt.grad.cpu().numpy(), gt_ref) <= 1e-4


def test_autograd_finite_differences(olv):
    """<grad, delta> against fp64 central differences of L = <c, r(w, t)>
    computed with the oracle's fixed point, along random directions."""
    from olive_amd import autograd

    csr, t = _case(olv, 10, 7)
    g = olv.Graph(csr)
    dev = torch.device("cuda:0")
    w = torch.tensor(csr.weights, device=dev, requires_grad=True)
    tt = torch.tensor(t, device=dev, requires_grad=True)
    rng = np.random.default_rng(9)
    c = rng.standard_normal(csr.n)
    r = autograd.pagerank(g, w, tt, tol=1e-8)
    (r * torch.tensor(c, dtype=torch.float32, device=dev)).sum().backward()
    gw = w.grad.cpu().numpy().astype(np.float64)
    gt = tt.grad.cpu().numpy().astype(np.float64)
    w64, t64 = csr.weights.astype(np.float64), t.astype(np.float64)
    L = lambda ww, t_: float(c @ oracle.pagerank_weighted(csr.row_ptr, csr.col, ww, t_, tol=1e-15)[0])
    for _ in range(3):
        dw, dt = rng.standard_normal(csr.m), rng.standard_normal(csr.n)
        h = 1e-5
        fd = (L(w64 + h * dw, t64 + h * dt) - L(w64 - h * dw, t64 - h * dt)) / (2 * h)
        an = gw @ dw + gt @ dt
        assert abs(fd - an) <= 1e-4 * max(abs(fd), 1e-3), (fd, an)


def test_autograd_reduces_to_plain_pagerank(olv):
    """w = 1/outdeg, t = 1/N: the forward is testPageRank's update run to
    convergence."""
    from olive_amd import autograd

    csr = olv.gen_rmat(10, 16, seed=3)
    g = olv.Graph(csr)
    od = csr.outdeg().astype(np.float32)
    w = np.repeat(1.0 / np.maximum(od, 1), csr.outdeg()).astype(np.float32)
    t = np.full(csr.n, 1.0 / csr.n, np.float32)
    dev = torch.device("cuda:0")
    r = autograd.pagerank(g, torch.tensor(w, device=dev), torch.tensor(t, device=dev), tol=1e-9, max_iters=500)
    ref, _, _ = oracle.pagerank(csr.row_ptr, csr.col, 1000, epsilon=1e-13)
    rr = r.cpu().numpy()
    assert float(np.max(np.abs(rr - ref) / ref)) <= 1e-4


def test_autograd_rejects_bad_input(olv):
    from olive_amd import autograd

    csr, t = _case(olv, 8, 1)
    g = olv.Graph(csr)
    dev = torch.device("cuda:0")
    with pytest.raises(TypeError):
        autograd.pagerank(g, torch.tensor(csr.weights), torch.tensor(t, device=dev))
    with pytest.raises(ValueError):
        autograd.pagerank(g, torch.tensor(csr.weights[:-1], device=dev), torch.tensor(t, device=dev))


def test_autograd_forward_mode(olv):
    """torch.autograd.forward_ad through the wrapper: the tangent of r for
    tangents of w and t against the dense fp64 tangent solve, and a tangent
    of t alone (w's tangent zero)."""
    import torch.autograd.forward_ad as fwAD

    from olive_amd import autograd

    csr, t = _case(olv, 11, 25)
    n = csr.n
    rng = np.random.default_rng(3)
    g = olv.Graph(csr)
    dw = rng.standard_normal(csr.m).astype(np.float32)
    dt = rng.standard_normal(n).astype(np.float32)
    src = np.repeat(np.arange(n), csr.outdeg())
    dst = csr.col.astype(np.int64)
    P = np.zeros((n, n))
    np.add.at(P, (src, dst), csr.weights.astype(np.float64))
    dP = np.zeros((n, n))
    np.add.at(dP, (src, dst), dw.astype(np.float64))
    M = np.eye(n) - 0.85 * P.T
    r_ref = np.linalg.solve(M, 0.15 * t.astype(np.float64))
    w_t = torch.tensor(csr.weights, device="cuda")
    t_t = torch.tensor(t, device="cuda")
    with fwAD.dual_level():
        r = autograd.pagerank(g, fwAD.make_dual(w_t, torch.tensor(dw, device="cuda")),
                              fwAD.make_dual(t_t, torch.tensor(dt, device="cuda")), tol=1e-8)
        dr = fwAD.unpack_dual(r).tangent.cpu().numpy()
    ref = np.linalg.solve(M, 0.85 * dP.T @ r_ref + 0.15 * dt.astype(np.float64))
    assert _rel(dr, ref) <= 1e-4
    with fwAD.dual_level():
        r = autograd.pagerank(g, w_t, fwAD.make_dual(t_t, torch.tensor(dt, device="cuda")), tol=1e-8)
        dr = fwAD.unpack_dual(r).tangent.cpu().numpy()
    assert _rel(dr, np.linalg.solve(M, 0.15 * dt.astype(np.float64))) <= 1e-4


def test_autograd_weights_updated_in_place(olv):
    """A training-style loop: the weights tensor is updated in place between
    steps (its version counter moves), so every step sees the new weights;
    steps with unchanged weights reuse the library's slot-order copy and
    give the same gradients as a fresh tensor."""
    from olive_amd import autograd

    csr, t = _case(olv, 11, 31)
    g = olv.Graph(csr)
    dev = torch.device("cuda:0")
    w = torch.tensor(csr.weights, device=dev, requires_grad=True)
    tt = torch.tensor(t, device=dev)
    c = torch.tensor(np.random.default_rng(2).standard_normal(csr.n).astype(np.float32), device=dev)
    for step in range(3):
        grads = []
        for rep in range(2):  # the second pass finds the weights unchanged
            w.grad = None
            r = autograd.pagerank(g, w, tt, tol=1e-8)
            (r * c).sum().backward()
            grads.append((r.detach().clone(), w.grad.clone()))
        # the same step with a fresh copy of the weights (new address)
        w2 = w.detach().clone().requires_grad_(True)
        r2 = autograd.pagerank(g, w2, tt, tol=1e-8)
        (r2 * c).sum().backward()
        for r, gw in grads:
            assert torch.equal(r, r2.detach()), step
            assert torch.equal(gw, w2.grad), step
        with torch.no_grad():
            w.mul_(1.0 + 0.1 * torch.rand_like(w))
