"""Differentiable sparse solve: ``x = A^{-1} b`` as a torch autograd function
on top of the GPU Krylov path, with implicit-differentiation backward
(the role of the reference's SPQR layer, sparse_bridge/SPQR_layer.py:22-58 and
bridge.cpp:151-185, which does this with a CPU QR factorisation).

    lam = A^{-T} dL/dx          (one more solve; A itself when A is symmetric)
    dL/db = lam
    dL/dA_ij = -lam_i x_j       (on the pattern, in the caller's entry order)

The matrix is stored in FP32 (the library's matrix format): ``values`` in
float32, or in float64 when every value is exactly representable in FP32;
other float64 values raise instead of being rounded silently, since the
solution and the gradients would then belong to a different matrix.
``b`` in float32 runs the FP32 solvers (relative residual ``tol``, default
1e-6); ``b`` in float64 runs FP64 iterative refinement around them (default
1e-10): x is then accurate to FP64 for the stored matrix, and so are the
gradients. A solve that does not converge raises.
"""
from __future__ import annotations

import torch

from . import load_ext

_C = load_ext()

__all__ = ["sparse_solve"]


def _fp32_values(values):
    """values as float32; float64 values must round-trip through FP32 exactly."""
    v = values.detach()
    if v.dtype == torch.float32:
        return v.contiguous()
    if v.dtype != torch.float64:
        raise TypeError(f"sparse_solve: values must be float32 or float64, got {v.dtype}")
    v32 = v.float()
    if not torch.equal(v32.double(), v):
        raise ValueError("sparse_solve: the matrix is stored in FP32 and these float64 values are not exactly "
                         "representable in FP32; pass values.float() (or values.float().double()) to solve the "
                         "FP32-rounded matrix")
    return v32.contiguous()


def _matrix(rows, cols, values, n):
    keys = (rows.to(torch.int64) << 32) | cols.to(torch.int64)
    e = torch.empty(0)
    return _C.csr_tensor(keys.contiguous(), _fp32_values(values), e, e, e, n, n)


def _solve(A, b, method, tol, maxiter):
    if b.dtype == torch.float64:
        x = torch.zeros_like(b)
        info = A.solve_refined(b.contiguous(), x, tol=tol, solver=method, max_outer=50)
    else:
        x = torch.zeros_like(b, dtype=torch.float32)
        fn = A.cgd if method == "cg" else A.bicgstab
        info = fn(b.float().contiguous(), x, tol=tol, maxiter=maxiter)
    if info["status"] != "converged":
        raise RuntimeError(f"sparse_solve: {method} did not converge ({info})")
    return x


class _SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, rows, cols, n, method, tol, maxiter):
        A = _matrix(rows, cols, values, n)
        x = _solve(A, b.detach(), method, tol, maxiter)
        ctx.save_for_backward(rows, cols, values, x)
        ctx.A, ctx.n, ctx.method, ctx.tol, ctx.maxiter = A, n, method, tol, maxiter
        return x

    @staticmethod
    def backward(ctx, g):
        rows, cols, values, x = ctx.saved_tensors
        # symmetric systems (CG) reuse A; otherwise solve with the transpose
        At = ctx.A if ctx.method == "cg" else _matrix(cols, rows, values, ctx.n)
        lam = _solve(At, g.to(x.dtype), ctx.method, ctx.tol, ctx.maxiter)
        grad_v = grad_b = None
        if ctx.needs_input_grad[0]:
            grad_v = -(lam[rows.long()] * x[cols.long()]).to(values.dtype)
        if ctx.needs_input_grad[1]:
            grad_b = lam
        return grad_v, grad_b, None, None, None, None, None, None


def sparse_solve(rows, cols, values, b, method="cg", tol=None, maxiter=10000):
    """x with A x = b, A the n x n matrix with entries values[k] at (rows[k],
    cols[k]) (duplicates added), n = b.numel(). Differentiable in `values` and
    `b`. method: "cg" (A symmetric positive definite) or "bicgstab"."""
    if method not in ("cg", "bicgstab"):
        raise ValueError("method must be 'cg' or 'bicgstab'")
    if b.dim() != 1:
        raise ValueError("b must be a vector")
    if tol is None:
        tol = 1e-10 if b.dtype == torch.float64 else 1e-6
    _fp32_values(values)  # refuse values the FP32 matrix cannot hold before any work
    return _SparseSolve.apply(values, b, rows, cols, b.numel(), method, float(tol), int(maxiter))
