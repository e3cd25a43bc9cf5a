"""Oracle for the direct sparse solve (test infrastructure only: imported by
tests/, never by the product path).

The reference factors with Accelerate's sparse QR (sparse_bridge/bridge.cpp:
14-191, Apple's closed library, absent here) and checks itself against
scipy.sparse.linalg.splu with ||x - x_splu|| < 1e-8 (sparse_bridge/
benchmark.py:62-86). What the result must be is fixed by the mathematics: the
unique solution of A x = b for square nonsingular A, the least-squares
solution for tall full-rank A. This module states it densely in FP64 (numpy
LAPACK) and gives the reference's own benchmark matrix.
"""
from __future__ import annotations

import numpy as np


def dense_from_csc(col_starts, row_indices, values, m: int, n: int) -> np.ndarray:
    """Dense FP64 matrix of a CSC triple (duplicates added)."""
    A = np.zeros((m, n), dtype=np.float64)
    cs = np.asarray(col_starts, dtype=np.int64)
    cols = np.repeat(np.arange(n), np.diff(cs))
    np.add.at(A, (np.asarray(row_indices, dtype=np.int64), cols), np.asarray(values, dtype=np.float64))
    return A


def least_squares(col_starts, row_indices, values, m: int, n: int, B) -> np.ndarray:
    """argmin ||A x - b|| per column of B (the solution for square nonsingular A)."""
    A = dense_from_csc(col_starts, row_indices, values, m, n)
    B = np.asarray(B, dtype=np.float64)
    if m == n:
        return np.linalg.solve(A, B)
    return np.linalg.lstsq(A, B, rcond=None)[0]


def laplacian_1d_shifted_csc(n: int, shift: float = 0.1):
    """sparse_bridge/benchmark.py:20-36 generate_data(n): 2 on the diagonal,
    -1 next to it, plus shift * I, as CSC (col_starts, row_indices, values)."""
    import scipy.sparse as sp

    A = (sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], shape=(n, n), format="csc")
         + shift * sp.eye(n, format="csc")).tocsc()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64)


def bandwidths(col_starts, row_indices, n: int, prow=None, pcol=None):
    """(kl, ku) of the pattern with rows / columns renumbered by prow / pcol."""
    cs = np.asarray(col_starts, dtype=np.int64)
    rows = np.asarray(row_indices, dtype=np.int64)
    cols = np.repeat(np.arange(n), np.diff(cs))
    if prow is not None:
        rows = np.asarray(prow)[rows]
    if pcol is not None:
        cols = np.asarray(pcol)[cols]
    d = cols - rows
    return int(max(0, -d.min(initial=0))), int(max(0, d.max(initial=0)))
