"""Weighted personalized PageRank behind ``torch.autograd.Function``.

The reference has no differentiable PageRank; its PageRank is the uniform
teleport pull iteration of include/velographx/incremental/pagerank.hpp:277-321.
This module adds the weighted, personalized form and its adjoint:

    P_uv = w_uv / W_u            (W_u = sum of u's out-weights)
    X    = (1 - d) S + d P^T X   (S, X: n x K teleport / rank columns)

Forward and backward run as hand-written gfx950 kernels in ``libvgx.so``
(``vgx_ppr_forward`` / ``vgx_ppr_backward``) on torch's current stream; torch
only holds the device tensors.  The backward pass solves the adjoint system
L = C + d P L with the same number of sweeps and returns

    dLoss/dS   = (1 - d) L
    dLoss/dw_uv = (G_uv - sum_v' P_uv' G_uv') / W_u,  G_uv = d <X_u, L_v>

which is the exact gradient of the fixed point (the truncation of the
iteration at ``iterations`` sweeps is not differentiated).

Edge weights are indexed by the arcs of the graph's out-CSR
(``Graph.export_csr()``: rows by source id, neighbours ascending).  The
graph must not change between forward and backward.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _velographx as _vx

_MAX_COLS = 32  # columns per kernel call (one half-wave lane per column)


def arc_list(graph) -> tuple[np.ndarray, np.ndarray]:
    """(src, dst) of every arc in the order edge weights are indexed."""
    off, adj = graph.export_csr()
    src = np.repeat(np.arange(len(off) - 1, dtype=np.uint32), np.diff(off).astype(np.int64))
    return src, np.asarray(adj, dtype=np.uint32)


def _check_device_tensor(name: str, t: torch.Tensor) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on a GPU (no CPU path exists)")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32")


class PersonalizedPageRank(torch.autograd.Function):
    """X = PersonalizedPageRank.apply(graph, weights, seeds, damping, iterations)."""

    @staticmethod
    def forward(ctx, graph, weights, seeds, damping, iterations):
        _check_device_tensor("weights", weights)
        _check_device_tensor("seeds", seeds)
        n = graph.vertex_count
        if seeds.dim() != 2 or seeds.shape[0] != n:
            raise ValueError(f"seeds must have shape (vertex_count={n}, K)")
        if weights.dim() != 1 or weights.shape[0] != graph.edge_count_directed:
            raise ValueError(f"weights must have shape (arcs={graph.edge_count_directed},)")
        w = weights.detach().contiguous()
        s = seeds.detach()
        K = s.shape[1]
        x = torch.empty((n, K), dtype=torch.float32, device=s.device)
        stream = torch.cuda.current_stream(s.device).cuda_stream
        for j in range(0, K, _MAX_COLS):
            k = min(_MAX_COLS, K - j)
            sj = s[:, j : j + k].contiguous()
            xj = torch.empty_like(sj)
            _vx.ppr_forward(graph, w.data_ptr(), sj.data_ptr(), xj.data_ptr(), k, float(damping),
                            int(iterations), stream)
            x[:, j : j + k] = xj
        ctx.graph = graph
        ctx.version = graph.version
        ctx.damping = float(damping)
        ctx.iterations = int(iterations)
        ctx.save_for_backward(w, x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        graph = ctx.graph
        if graph.version != ctx.version:
            raise RuntimeError("graph changed between forward and backward")
        w, x = ctx.saved_tensors
        need_w, need_s = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        c = grad_x.detach().to(torch.float32)
        n, K = x.shape
        grad_w = torch.zeros_like(w) if need_w else None
        grad_s = torch.empty((n, K), dtype=torch.float32, device=x.device) if need_s else None
        stream = torch.cuda.current_stream(x.device).cuda_stream
        for j in range(0, K, _MAX_COLS):
            k = min(_MAX_COLS, K - j)
            cj = c[:, j : j + k].contiguous()
            xj = x[:, j : j + k].contiguous()
            lam = torch.empty_like(cj)
            gs = torch.empty_like(cj) if need_s else None
            gw = torch.empty_like(w) if need_w else None
            _vx.ppr_backward(graph, w.data_ptr(), xj.data_ptr(), cj.data_ptr(), lam.data_ptr(),
                             gs.data_ptr() if gs is not None else 0, gw.data_ptr() if gw is not None else 0,
                             k, ctx.damping, ctx.iterations, stream)
            if need_s:
                grad_s[:, j : j + k] = gs
            if need_w:
                grad_w += gw
        return None, grad_w, grad_s, None, None


def personalized_pagerank(graph, weights: torch.Tensor, seeds: torch.Tensor, damping: float = 0.85,
                          iterations: int = 100) -> torch.Tensor:
    """Differentiable weighted personalized PageRank columns (n x K)."""
    if not 0.0 <= damping <= 1.0:
        raise ValueError("damping must be in [0, 1]")
    if iterations < 0:
        raise ValueError("iterations must be non-negative")
    return PersonalizedPageRank.apply(graph, weights, seeds, damping, iterations)
