"""FSAI / edge-MLP oracle (test infrastructure only).

The reference has no preconditioner; these restate the library's learned-FSAI
additions from their definitions so the GPU kernels can be checked:

* :func:`lower_pattern`, :func:`features` -- the lower-triangle pattern of A in
  CSR order and the 8 per-edge features (the row couplings s_low / s_all are pinned
  by the closed form of classical FSAI on stencil matrices, whose lower neighbours
  are not coupled to each other: g_ii sqrt(a_ii) = 1 / sqrt(1 - s_low));
* :func:`classical_fsai` -- per row i: A[J,J] y = e_i, g = y / sqrt(y_i)
  (Kolotilina-Yeremin FSAI); pinned by its defining property (G A)[i, J] =
  e_i / sqrt(y_i) and by the diagonal case G = D^{-1/2};
* :func:`probe_loss_grad` -- L = sum_w ||G A G^T w - w||^2 and dL/dG on the
  pattern; pinned by central finite differences;
* :func:`rownorm` / :func:`rownorm_grad` -- the stated-5 network's row
  normalisation diag(G A G^T) = 1 and its chain rule; pinned by classical
  FSAI (a row shape from classical FSAI normalises back to classical FSAI:
  it has unit diag(G A G^T) by construction) and by central finite differences;
* :func:`mlp_forward` / :func:`mlp_backward` -- 8 -> 64 -> 64 -> 1 ReLU network
  in float64, optionally rounding the matrix-core operands to BF16 exactly where
  the kernels do; pinned against torch autograd (float64) with rounding off.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .spmv import round_bf16

KIN, KH = 8, 64
OW1, OB1 = 0, 512
OW2 = OB1 + KH
OB2 = OW2 + KH * KH
OW3 = OB2 + KH
OB3 = OW3 + KH
NPARAMS = OB3 + 1


def lower_pattern(row_ptr, col_ind):
    """Edges (i, j <= i) of A's pattern in CSR order: (rows, cols, src positions)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col = np.asarray(col_ind, dtype=np.int64)
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    keep = col <= rows
    return rows[keep], col[keep], np.nonzero(keep)[0]


def diag_of(row_ptr, col_ind, vals):
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    d = np.zeros(n)
    m = np.asarray(col_ind, dtype=np.int64) == rows
    np.add.at(d, rows[m], np.asarray(vals, dtype=np.float64)[m])
    return d


def features(row_ptr, col_ind, vals, feature_set="coupled8"):
    """(feat[ne, 8], colscale[ne]) for the lower pattern (float64).

    "coupled8": the 8 features with the row couplings s_low, s_all, t;
    "stated5": the edge's own five quantities (a_ij, a_ii, a_jj, row nnz,
    row sum |a|) made scale-free -- the first five columns of "coupled8" --
    and zeros in the last three."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    n = row_ptr.size - 1
    r, c, src = lower_pattern(row_ptr, col_ind)
    d = diag_of(row_ptr, col_ind, vals)
    rows_all = np.repeat(np.arange(n), np.diff(row_ptr))
    asum = np.zeros(n)
    np.add.at(asum, rows_all, np.abs(vals))
    f = np.zeros((r.size, 8))
    f[:, 0] = vals[src] / np.sqrt(d[r] * d[c])
    f[:, 1] = np.log(d[c] / d[r])
    f[:, 2] = np.diff(row_ptr)[r] / 8.0
    f[:, 3] = asum[r] / d[r]
    f[:, 4] = (r == c).astype(np.float64)
    cols_all = np.asarray(col_ind, dtype=np.int64)
    off = cols_all != rows_all
    q = np.where(off, vals ** 2 / (d[rows_all] * d[cols_all]), 0.0)
    s_all = np.zeros(n)
    np.add.at(s_all, rows_all, q)
    s_low = np.zeros(n)
    np.add.at(s_low, rows_all, np.where(cols_all < rows_all, q, 0.0))
    f[:, 5] = s_low[r]
    f[:, 6] = s_all[r]
    f[:, 7] = 1.0 / np.sqrt(np.maximum(1.0 - s_low[r], 1.0 / 64.0))
    if feature_set == "stated5":
        f[:, 5:] = 0.0
    elif feature_set != "coupled8":
        raise ValueError(feature_set)
    return f, 1.0 / np.sqrt(d[c])


def rownorm(A, rows, cols, gt):
    """(G, s): G_i = gt_i / sqrt(N_i), N_i = gt_i A[J,J] gt_i^T, so that
    diag(G A G^T) = 1 (s_i = N_i^{-1/2}; rows with N_i <= 0 keep gt and get
    s_i = 0). A dense quadratic form per row -- the definition itself."""
    A = sp.csr_matrix(A, dtype=np.float64)
    n = A.shape[0]
    Gt = G_matrix(n, rows, cols, gt)
    N = np.asarray((Gt @ A).multiply(Gt).sum(axis=1)).ravel()  # (G~ A G~^T)_ii
    ok = N > 0
    s = np.where(ok, 1.0 / np.sqrt(np.where(ok, N, 1.0)), 0.0)
    G = np.asarray(gt, dtype=np.float64) * np.where(ok, s, 1.0)[rows]
    return G, s


def rownorm_grad(A, rows, cols, gt, gG):
    """dL/dgt from dL/dG through rownorm (chain rule of G = gt / sqrt(N)):
    dL/dgt_p = s gG_p - s^3 v_p sum_q gG_q gt_q with v = A[J,J] gt."""
    A = sp.csr_matrix(A, dtype=np.float64)
    n = A.shape[0]
    gt = np.asarray(gt, dtype=np.float64)
    gG = np.asarray(gG, dtype=np.float64)
    _, s = rownorm(A, rows, cols, gt)
    Gt = G_matrix(n, rows, cols, gt)
    V = (Gt @ A).tocsr()  # V[i, j] = (gt_i A)_j; v_p = V[i, J_p] (A symmetric)
    v = np.asarray(V[rows, cols]).ravel()
    c = np.bincount(rows, weights=gG * gt, minlength=n)
    ok = s[rows] > 0
    return np.where(ok, s[rows] * gG - s[rows] ** 3 * v * c[rows], gG)


def classical_fsai(A: sp.csr_matrix):
    """Classical FSAI values on the lower pattern of A (float64)."""
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sort_indices()
    r, c, _ = lower_pattern(A.indptr, A.indices)
    g = np.zeros(r.size)
    n = A.shape[0]
    starts = np.searchsorted(r, np.arange(n + 1))
    Ad = A.tolil()
    for i in range(n):
        J = c[starts[i]:starts[i + 1]]
        S = np.array(Ad[J][:, J].todense())
        e = np.zeros(J.size)
        e[-1] = 1.0
        y = np.linalg.solve(S, e)
        g[starts[i]:starts[i + 1]] = y / np.sqrt(y[-1])
    return g


def G_matrix(n, rows, cols, g):
    return sp.csr_matrix((np.asarray(g, dtype=np.float64), (rows, cols)), shape=(n, n))


def probe_loss_grad(A, rows, cols, g, W):
    """L = sum_k ||G A G^T w_k - w_k||^2 and dL/dg (per pattern edge)."""
    n = A.shape[0]
    G = G_matrix(n, rows, cols, g)
    loss = 0.0
    grad = np.zeros(len(g))
    for w in np.atleast_2d(W):
        w = np.asarray(w, dtype=np.float64)
        u = G.T @ w
        y = A @ u
        e = G @ y - w
        h = A @ (G.T @ e)
        loss += float(e @ e)
        grad += 2.0 * (e[rows] * y[cols] + w[rows] * h[cols])
    return loss, grad


def _bf16(a, on):
    """FP32 rounding, then BF16 round-to-nearest-even (oracle.spmv.round_bf16), as float64."""
    if not on:
        return np.asarray(a, dtype=np.float64)
    return round_bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OW1:OB1].reshape(KH, KIN), P[OB1:OW2], P[OW2:OB2].reshape(KH, KH), P[OB2:OW3], P[OW3:OB3], P[OB3])


def mlp_forward(P, F, bf16=True):
    """out[e] and the activations (H1, H2) -- matrix-core operands rounded to BF16 when asked."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    H1 = np.maximum(_bf16(F, bf16) @ _bf16(W1, bf16).T + b1, 0.0)
    H2 = np.maximum(_bf16(H1, bf16) @ _bf16(W2, bf16).T + b2, 0.0)
    out = H2 @ w3 + b3
    return out, H1, H2


def mlp_backward(P, F, go, bf16=True):
    """d(sum_e go[e] out[e]) / dP, with the kernels' BF16 operand roundings."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    _, H1, H2 = mlp_forward(P, F, bf16)
    go = np.asarray(go, dtype=np.float64)
    dH2 = np.where(H2 > 0, go[:, None] * w3[None, :], 0.0)
    g = np.zeros(NPARAMS)
    g[OW3:OB3] = go @ H2
    g[OB3] = go.sum()
    g[OB2:OW3] = dH2.sum(0)
    g[OW2:OB2] = (_bf16(dH2, bf16).T @ _bf16(H1, bf16)).ravel()
    dH1 = np.where(H1 > 0, _bf16(dH2, bf16) @ _bf16(W2, bf16), 0.0)
    g[OB1:OW2] = dH1.sum(0)
    g[OW1:OB1] = (_bf16(dH1, bf16).T @ _bf16(F, bf16)).ravel()
    return g
