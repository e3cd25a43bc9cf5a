"""Direct sparse solve on the GPU: drop-in for the reference's sparse QR layer.

``SparseQRSolver(col_starts, row_indices, values, n_rows, n_cols)`` mirrors
sparse_bridge/bridge.cpp:14-191 (CSC with int64 indices and float64 values;
``solve(B)`` for a vector or a matrix of right-hand sides; ``get_grad_a_helper``).
The factorisation is a banded Householder QR in FP64 on the GPU after a
bandwidth-reducing ordering (reverse Cuthill-McKee); matrices whose ordering
leaves kl + ku + 1 > 256 or kl > 127 are refused (tall ones need
kl >= n_rows - n_cols). Tensors may be CPU or ROCm tensors; X
comes back on B's device.

``SparseSolve.apply(solver, A_sparse, b)`` mirrors SparseSolve in
sparse_bridge/SPQR_layer.py:12-58: x = A^{-1} b with the implicit-function
backward (dL/db = A^{-T} dL/dx through a solver of A^T, dL/dA_ij = -dL/db_i x_j
on A's pattern). Unlike the reference, the gradient for A is laid out in
A_sparse's own (coalesced COO) entry order.
"""
from __future__ import annotations

import torch

from . import load_ext

_C = load_ext()
SparseQRSolver = _C.SparseQRSolver

__all__ = ["SparseQRSolver", "SparseSolve", "solver_from_sparse"]


def solver_from_sparse(A_sparse: torch.Tensor) -> "SparseQRSolver":
    """A solver for a torch sparse (COO or CSC) float64 matrix."""
    csc = A_sparse.coalesce().to_sparse_csc() if A_sparse.layout == torch.sparse_coo else A_sparse.to_sparse_csc()
    return SparseQRSolver(csc.ccol_indices().to(torch.int64).contiguous(), csc.row_indices().to(torch.int64).contiguous(),
                          csc.values().to(torch.float64).contiguous(), A_sparse.shape[0], A_sparse.shape[1])


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, A_sparse, b):
        x = solver.solve(b.detach())
        ctx.save_for_backward(A_sparse, x)
        ctx.solver = solver
        return x

    @staticmethod
    def backward(ctx, grad_output):
        A_sparse, x = ctx.saved_tensors
        if A_sparse.shape[0] != A_sparse.shape[1]:
            raise RuntimeError("SparseSolve backward needs a square matrix")
        t_solver = solver_from_sparse(A_sparse.detach().t().coalesce())
        g = grad_output.contiguous()
        grad_b = t_solver.solve(g)
        grad_a = None
        if ctx.needs_input_grad[1]:
            if grad_b.dim() > 1:
                grad_a = None  # as the reference: no batched gradient for A
            else:
                A = A_sparse.detach().coalesce()
                rows, cols = A.indices()
                vals = -(grad_b[rows.to(grad_b.device)] * x[cols.to(x.device)])
                grad_a = torch.sparse_coo_tensor(A.indices(), vals.to(A.values().device), A.size())
        return None, grad_a, grad_b
