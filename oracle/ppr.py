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
tests/test_oracle.py by what the mathematics fixes: d = 0 gives X = S; a
directed cycle has a closed form that a transposed P fails; the solve and
the Jacobi limit agree; the adjoint gradients equal FP64 central finite
differences of the solve.  Arcs are given as (src, dst, w) arrays; their
order is the caller's.
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
