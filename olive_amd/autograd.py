"""torch.autograd wrapper of the weighted PageRank fixed point and its
hand-written adjoint (olv_pagerank_weighted_f32 / olv_pagerank_vjp_f32).

Forward:  r = d * P^T r + (1 - d) * t,  P[u, v] = w[e] for edge e = (u, v),
          iterated from r = t until ||r_k+1 - r_k||_1 < tol.
Backward: lambda = gbar + d * P lambda (the adjoint fixed point, solved on the
          out-CSR), then dL/dt = (1 - d) * lambda and
          dL/dw[e] = d * r[u] * lambda[v] for edge e = (u, v).
Forward mode (torch.autograd.forward_ad): the tangent
          dr = d * P^T dr + d * dP^T r + (1 - d) * dt
          (olv_pagerank_jvp_f32), for tangents of w and/or t.

The gradient is taken with respect to the weights as given (no chain rule
through a per-source normalisation the caller may have applied).  With
w = 1/outdeg and t = 1/N the forward is the update of testPageRank.cu:83-88.

Tensors are float32 on the GPU; the library runs on torch's current stream
for the duration of each call.  There is no CPU path: without the HIP library
the calls raise.
"""
from __future__ import annotations

import contextlib
import ctypes
import weakref

from . import Graph, check, lib
from ._lib import OLV_W_UNCHANGED

__all__ = ["PageRankFunction", "pagerank"]


@contextlib.contextmanager
def _torch_stream(device):
    import torch

    prev = lib().olv_get_stream()
    check(lib().olv_set_stream(ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)), "olv_set_stream")
    try:
        yield
    finally:
        lib().olv_set_stream(ctypes.c_void_p(prev))


def _check_vec(x, size, name):
    import torch

    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if x.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {x.dtype}")
    if x.dim() != 1 or x.numel() != size:
        raise ValueError(f"{name} must have shape ({size},), got {tuple(x.shape)}")
    return x.contiguous()


def _weights_flag(graph, w):
    """OLV_W_UNCHANGED when `w` is the tensor object the previous weighted
    call on this graph received, at the same address, and torch's version
    counter shows no in-place change since: the library may then reuse the
    copy of the weights it already put in its solvers' order.  Writes
    that bypass torch (raw pointers) are not seen, as with any autograd
    version check."""
    last = getattr(graph, "_w_seen", None)
    same = last is not None and last[0]() is w and last[1] == w._version and last[2] == w.data_ptr()
    graph._w_seen = (weakref.ref(w), w._version, w.data_ptr())
    return OLV_W_UNCHANGED if same else 0


class PageRankFunction:
    """`PageRankFunction.apply(graph, w, t, damping, tol, max_iters)` -> r.

    graph: olive_amd.Graph (built with its transpose); w: (m,) edge weights in
    CSR edge order; t: (n,) teleport vector.  Differentiable in w and t."""

    _fn = None

    @classmethod
    def apply(cls, graph, w, t, damping=0.85, tol=1e-6, max_iters=1000):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(graph, w, t, float(damping), float(tol), int(max_iters))


def _make_function():
    import torch

    class _PageRank(torch.autograd.Function):
        @staticmethod
        def forward(ctx, graph: Graph, w, t, damping, tol, max_iters):
            w = _check_vec(w, graph.m, "w")
            t = _check_vec(t, graph.n, "t")
            r = torch.empty(graph.n, dtype=torch.float32, device=t.device)
            with _torch_stream(t.device):
                check(lib().olv_pagerank_weighted_f32(graph.h, w.data_ptr(), t.data_ptr(), damping, tol, max_iters,
                                                      r.data_ptr(), _weights_flag(graph, w), None),
                      "olv_pagerank_weighted_f32")
            ctx.graph = graph
            ctx.hyper = (damping, tol, max_iters)
            ctx.save_for_backward(w, r)
            ctx.save_for_forward(w, r)
            return r

        @staticmethod
        def backward(ctx, gbar):
            w, r = ctx.saved_tensors
            graph = ctx.graph
            damping, tol, max_iters = ctx.hyper
            gbar = gbar.detach().to(torch.float32).contiguous()
            gw = torch.empty_like(w)
            gt = torch.empty_like(r)
            lam = torch.empty_like(r)
            with _torch_stream(r.device):
                check(lib().olv_pagerank_vjp_f32(graph.h, w.data_ptr(), r.data_ptr(), gbar.data_ptr(), damping, tol,
                                                 max_iters, gw.data_ptr(), gt.data_ptr(), lam.data_ptr(),
                                                 _weights_flag(graph, w), None),
                      "olv_pagerank_vjp_f32")
            return None, gw, gt, None, None, None

        @staticmethod
        def jvp(ctx, _graph_t, dw, dt, *_):
            w, r = ctx.saved_tensors
            graph = ctx.graph
            damping, tol, max_iters = ctx.hyper
            dw = torch.zeros_like(w) if dw is None else dw.detach().to(torch.float32).contiguous()
            dt = torch.zeros_like(r) if dt is None else dt.detach().to(torch.float32).contiguous()
            dr = torch.empty_like(r)
            with _torch_stream(r.device):
                check(lib().olv_pagerank_jvp_f32(graph.h, w.data_ptr(), r.data_ptr(), dw.data_ptr(), dt.data_ptr(),
                                                 damping, tol, max_iters, dr.data_ptr(), _weights_flag(graph, w),
                                                 None),
                      "olv_pagerank_jvp_f32")
            return dr

    return _PageRank


def pagerank(graph: Graph, w, t, damping=0.85, tol=1e-6, max_iters=1000):
    """Differentiable weighted PageRank (see the module docstring)."""
    return PageRankFunction.apply(graph, w, t, damping, tol, max_iters)
