"""FP64 oracle for weighted personalized PageRank -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s checks import this module, as the checker.

The reference has no weighted or personalized PageRank: its PageRank
(include/velographx/incremental/pagerank.hpp:277-321) is the pull iteration
with uniform teleport ``(1-d)/n`` and dangling mass spread uniformly.  The
differentiable form added here keeps the damping/teleport split and changes
the two things personalization changes:

    P_uv = w_uv / W_u,   W_u = sum_v w_uv   (rows with W_u = 0 send no mass)
    X    = (1 - d) S + d P^T X              (S: n x K teleport columns)

and, for a loss with dLoss/dX = C,

    L     = C + d P L                       (adjoint system)
    dS    = (1 - d) L
    G_uv  = d <X_u, L_v>                    (gradient w.r.t. P_uv)
    dw_uv = (G_uv - sum_v' P_uv' G_uv') / W_u

Nothing in the reference pins these, so every function here is pinned in
tests/test_oracle_ppr.py by what the mathematics fixes: d = 0 gives X = S; a
directed cycle has a closed form that a transposed P fails; the solve and
the Jacobi limit agree; the adjoint gradients equal FP64 central finite
differences of the solve; the C sweeps and the rank-one finite differences
used at 2^20 vertices equal the numpy sweeps and the dense-solve finite
differences.  Arcs are given as (src, dst, w) arrays; their order is the
caller's.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def transition(n: int, src, dst, w):
    """Row-normalised transition matrix P (CSR, float64) and row sums W."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.float64)
    W = np.bincount(src, weights=w, minlength=n).astype(np.float64)
    scale = np.divide(1.0, W, out=np.zeros_like(W), where=W > 0)
    P = sp.csr_matrix((w * scale[src], (src, dst)), shape=(n, n))
    return P, W


def forward(n: int, src, dst, w, S, d: float, iterations: int | None = None, dense: bool = False):
    """X for teleport columns S.  ``iterations=None`` solves the system
    exactly (dense LU if ``dense`` else sparse LU); otherwise X_0 = (1-d) S
    followed by ``iterations`` Jacobi sweeps."""
    P, _ = transition(n, src, dst, w)
    S = np.asarray(S, np.float64).reshape(n, -1)
    if iterations is None:
        if dense:
            A = np.eye(n) - d * P.T.toarray()
            return np.linalg.solve(A, (1.0 - d) * S)
        A = (sp.identity(n, format="csc") - d * P.T.tocsc()).tocsc()
        return np.asarray(spla.splu(A).solve((1.0 - d) * S))
    PT = P.T.tocsr()
    X = (1.0 - d) * S
    for _ in range(iterations):
        X = (1.0 - d) * S + d * (PT @ X)
    return X


def adjoint(n: int, src, dst, w, C, d: float, iterations: int | None = None, dense: bool = False):
    """L solving L = C + d P L (exactly, or by Jacobi sweeps from L_0 = C)."""
    P, _ = transition(n, src, dst, w)
    C = np.asarray(C, np.float64).reshape(n, -1)
    if iterations is None:
        if dense:
            return np.linalg.solve(np.eye(n) - d * P.toarray(), C)
        A = (sp.identity(n, format="csc") - d * P.tocsc()).tocsc()
        return np.asarray(spla.splu(A).solve(C))
    L = C.copy()
    for _ in range(iterations):
        L = C + d * (P @ L)
    return L


def gradients(n: int, src, dst, w, S, C, d: float, iterations: int | None = None, dense: bool = False):
    """(X, dLoss/dw per arc, dLoss/dS) for Loss = sum(C * X)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.float64)
    X = forward(n, src, dst, w, S, d, iterations, dense)
    L = adjoint(n, src, dst, w, C, d, iterations, dense)
    P, W = transition(n, src, dst, w)
    G = d * np.einsum("ek,ek->e", X[src], L[dst])
    p = w / np.where(W[src] > 0, W[src], 1.0)
    row_dot = np.bincount(src, weights=p * G, minlength=n)
    gw = np.where(W[src] > 0, (G - row_dot[src]) / np.where(W[src] > 0, W[src], 1.0), 0.0)
    return X, gw, (1.0 - d) * L


def loss(n: int, src, dst, w, S, C, d: float, dense: bool = True) -> float:
    """sum(C * X) with X from the exact solve."""
    return float(np.sum(np.asarray(C, np.float64) * forward(n, src, dst, w, S, d, None, dense)))


def finite_difference_w(n: int, src, dst, w, S, C, d: float, arcs, eps: float = 1e-6, dense: bool = True):
    """FP64 central differences of ``loss`` w.r.t. the weights of ``arcs``."""
    w = np.asarray(w, np.float64)
    out = np.empty(len(arcs))
    for i, a in enumerate(arcs):
        wp = w.copy()
        wm = w.copy()
        h = eps * max(1.0, abs(w[a]))
        wp[a] += h
        wm[a] -= h
        out[i] = (loss(n, src, dst, wp, S, C, d, dense) - loss(n, src, dst, wm, S, C, d, dense)) / (2 * h)
    return out


# ------------------------------------------------------- at 2^20 vertices
# The numpy sweeps and the LU solves above do not finish in test time at
# 2^20 vertices and 16M arcs; the C sweeps (orc_ppr_sweeps, OpenMP over rows)
# do, and the rank-one identity below turns three of their solves into exact
# perturbed losses.


def csr_rows(n: int, keys, idx, p):
    """Rows of arcs grouped by `keys` (stable within a key): (off, idx, p)."""
    keys = np.asarray(keys, np.int64)
    order = np.argsort(keys, kind="stable")
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(keys, minlength=n), out=off[1:])
    return off, np.ascontiguousarray(np.asarray(idx)[order], np.uint32), np.ascontiguousarray(
        np.asarray(p, np.float64)[order])


def sweeps_c(n: int, rows, B, d: float, iterations: int):
    """Y_0 = B, then Y = B + d * (row-sum of p * Y[idx]) `iterations` times
    (orc_ppr_sweeps): the forward over the in-rows with B = (1 - d) S, the
    adjoint over the out-rows with B = C."""
    import ctypes

    from . import oracle as orc

    off, idx, p = rows
    B = np.ascontiguousarray(np.asarray(B, np.float64).reshape(n, -1))
    K = B.shape[1]
    Y = np.empty_like(B)
    L = orc.lib()
    L.orc_ppr_sweeps.restype = ctypes.c_int
    L.orc_ppr_sweeps.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_uint32, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint64,
                                 ctypes.c_void_p]
    rc = L.orc_ppr_sweeps(n, off.ctypes.data, idx.ctypes.data if idx.size else None,
                          p.ctypes.data if p.size else None, K, B.ctypes.data, float(d), int(iterations),
                          Y.ctypes.data)
    if rc != 0:
        raise ValueError("at most 64 columns")
    return Y


def forward_c(n: int, src, dst, w, S, d: float, iterations: int):
    """X after `iterations` sweeps from X_0 = (1 - d) S (same as forward())."""
    src = np.asarray(src, np.int64)
    w = np.asarray(w, np.float64)
    W = np.bincount(src, weights=w, minlength=n)
    p = w / np.where(W[src] > 0, W[src], 1.0)
    return sweeps_c(n, csr_rows(n, dst, src, p), (1.0 - d) * np.asarray(S, np.float64).reshape(n, -1), d,
                    iterations)


def finite_difference_w_rank1(n: int, src, dst, w, S, C, d: float, arcs, eps: float = 1e-5,
                              iterations: int = 150):
    """FP64 central differences of Loss = sum(C * X) w.r.t. the weights of
    `arcs`, each perturbed loss evaluated exactly from unperturbed solves.

    A change of w_a, a = (u, v), changes row u of P only: P' = P + e_u r^T,
    so I - d P'^T = A - d r e_u^T with A = I - d P^T, a rank-one change.  By
    the Sherman-Morrison identity

        Loss' - Loss = d * sum_j r_j <X_u, L_j> / (1 - d * sum_j r_j y_j)

    where X = (1 - d) A^-1 S (forward), L = A^-T C (adjoint: L = C + d P L),
    y = A^-T e_u (adjoint with a one-hot right-hand side) and j runs over
    u's out-neighbours.  Each of Loss(w +- h e_a) - Loss(w) is therefore
    exact up to the solves' truncation (d^(iterations+1)); the central
    difference of the two is free of the cancellation of subtracting two
    whole losses.  Costs three solves (the y of all sampled sources in one
    multi-column solve) instead of two per arc."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.float64)
    S = np.asarray(S, np.float64).reshape(n, -1)
    C = np.asarray(C, np.float64).reshape(n, -1)
    W = np.bincount(src, weights=w, minlength=n)
    p = w / np.where(W[src] > 0, W[src], 1.0)
    out_rows = csr_rows(n, src, dst, p)
    X = sweeps_c(n, csr_rows(n, dst, src, p), (1.0 - d) * S, d, iterations)
    L = sweeps_c(n, out_rows, C, d, iterations)
    arcs = np.asarray(arcs, np.int64)
    us = np.unique(src[arcs])
    if us.size > 64:
        raise ValueError("at most 64 distinct sources")
    E = np.zeros((n, us.size))
    E[us, np.arange(us.size)] = 1.0
    Y = sweeps_c(n, out_rows, E, d, iterations)
    col = {int(u): i for i, u in enumerate(us)}
    # arcs of every row in arc order (a row's arcs are contiguous in out-CSR order)
    order = np.argsort(src, kind="stable")
    starts = np.searchsorted(src[order], np.arange(n + 1))
    out = np.empty(arcs.size)
    for i, a in enumerate(arcs):
        u = int(src[a])
        row = order[starts[u]:starts[u + 1]]
        wr = w[row]
        Wu = wr.sum()
        xl = L[dst[row]] @ X[u]          # <X_u, L_j> for every j in the row
        yj = Y[dst[row], col[u]]
        h = eps * max(1.0, abs(w[a]))
        deltas = []
        for sgn in (1.0, -1.0):
            wp = wr.copy()
            wp[row == a] += sgn * h
            r = wp / (Wu + sgn * h) - wr / Wu
            deltas.append(d * float(r @ xl) / (1.0 - d * float(r @ yj)))
        out[i] = (deltas[0] - deltas[1]) / (2.0 * h)
    return out, X, L
