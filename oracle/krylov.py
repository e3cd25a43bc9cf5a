"""Krylov-solver oracle (test infrastructure only).

Restatements of the reference's solvers with the stopping rule the library uses
(relative residual ``||r|| <= tol * ||b||`` after each full iteration, no
lag) and the breakdown rules listed below. Arithmetic is in ``dtype`` (FP64 by
default; FP32 to mirror the GPU's working precision).

* :func:`cg` -- textbook CG as in iterative/BiCGSTAB/cgd.metal:11-90 driven by
  iterative/BiCGSTAB/spmv.cpp:528-659 (``r=b-Ax0, p=r``; ``alpha=rr/pAp``;
  ``x+=alpha p``; ``r-=alpha Ap``; ``p=r+(rr_new/rr)p``). Breakdown when
  ``pAp <= 0`` or non-finite (the reference instead sets alpha=0 when
  ``|pAp| <= 1e-9``, cgd.metal:40).
* :func:`bicgstab` -- van der Vorst BiCGSTAB in the kernel order of
  iterative/BiCGSTAB/bicgstab.metal:12-326 + spmv.cpp:699-888 (``upp``,
  ``fused_spvmv``, ``f_upalpha_ups``, ``spmv_op``, ``f_dip``,
  ``f_upomega_upxr_laa``) with rho0 = <r0, r0> (the reference starts from
  ``1 + <b,b>``, spmv.cpp:724,747-752 -- a quirk not reproduced). Guards: omega = 0
  when <t,t> = 0 (bicgstab.metal:265-269, exact zero instead of < 1e-20);
  beta = 0 when omega = 0 (restart, bicgstab.metal:48-55); breakdown when
  <r0,v> = 0 or rho = 0 or a scalar is non-finite.
* :func:`richardson` -- ``x <- x + omega (b - A x)`` for a fixed count
  (iterative/richardson/spmv.cpp:376-495, ``weighted_add`` spmv.metal:50-63).
* :func:`scipy_cg` / :func:`scipy_bicgstab` -- the CPU baseline solves.
"""
from __future__ import annotations

import numpy as np

CONVERGED = 0
MAXITER = 1
BREAKDOWN = 2


def _mv(A, x, dtype):
    return (A @ x.astype(np.float64)).astype(dtype) if dtype != np.float64 else A @ x


def _as(A, dtype):
    return A.astype(dtype) if A.dtype != dtype else A


def true_relres(A, x, b) -> float:
    """||b - A x|| / ||b|| in FP64."""
    A64 = _as(A, np.float64)
    x = np.asarray(x, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(b - A64 @ x) / (nb if nb > 0 else 1.0))


def cg(A, b, tol=1e-6, maxiter=10000, x0=None, dtype=np.float64, M_apply=None):
    """(Preconditioned) CG. Returns (x, iterations, status, history of ||r||/||b||).

    ``M_apply(r) -> z`` applies a preconditioner M ~ A^{-1} (None = plain CG).
    """
    A = _as(A, dtype)
    b = np.asarray(b, dtype=dtype)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=dtype).copy()
    r = b - (A @ x) if x0 is not None else b.copy()
    z = r if M_apply is None else np.asarray(M_apply(r), dtype=dtype)
    p = z.copy()
    rz = dtype(np.dot(r, z))
    nb = float(np.sqrt(np.dot(b.astype(np.float64), b.astype(np.float64))))
    thresh = tol * nb
    hist = [float(np.linalg.norm(r)) / (nb if nb else 1.0)]
    if np.linalg.norm(r) <= thresh:
        return x, 0, CONVERGED, hist
    for k in range(1, maxiter + 1):
        q = A @ p
        pq = dtype(np.dot(p, q))
        if not np.isfinite(pq) or pq <= 0:
            return x, k - 1, BREAKDOWN, hist
        alpha = dtype(rz / pq)
        x += alpha * p
        r -= alpha * q
        rnorm = float(np.linalg.norm(r.astype(np.float64)))
        hist.append(rnorm / (nb if nb else 1.0))
        if rnorm <= thresh:
            return x, k, CONVERGED, hist
        z = r if M_apply is None else np.asarray(M_apply(r), dtype=dtype)
        rz_new = dtype(np.dot(r, z))
        beta = dtype(rz_new / rz)
        p = z + beta * p
        rz = rz_new
    return x, maxiter, MAXITER, hist


def bicgstab(A, b, tol=1e-6, maxiter=10000, x0=None, dtype=np.float64):
    """BiCGSTAB. Returns (x, iterations, status, history of ||r||/||b||)."""
    A = _as(A, dtype)
    b = np.asarray(b, dtype=dtype)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=dtype).copy()
    r = b - (A @ x) if x0 is not None else b.copy()
    r0 = r.copy()
    rho = dtype(np.dot(r0, r))
    p = r.copy()
    nb = float(np.linalg.norm(b.astype(np.float64)))
    thresh = tol * nb
    hist = [float(np.linalg.norm(r)) / (nb if nb else 1.0)]
    if np.linalg.norm(r) <= thresh:
        return x, 0, CONVERGED, hist
    for k in range(1, maxiter + 1):
        v = A @ p
        sigma = dtype(np.dot(r0, v))
        if sigma == 0 or not np.isfinite(sigma):
            return x, k - 1, BREAKDOWN, hist
        alpha = dtype(rho / sigma)
        s = r - alpha * v
        t = A @ s
        phi1 = dtype(np.dot(t, s))
        phi2 = dtype(np.dot(t, t))
        omega = dtype(phi1 / phi2) if phi2 != 0 else dtype(0)
        x += alpha * p + omega * s
        r = s - omega * t
        rho_new = dtype(np.dot(r0, r))
        rnorm = float(np.linalg.norm(r.astype(np.float64)))
        hist.append(rnorm / (nb if nb else 1.0))
        if rnorm <= thresh:
            return x, k, CONVERGED, hist
        if rho_new == 0 or not np.isfinite(rho_new) or not np.isfinite(omega):
            return x, k, BREAKDOWN, hist
        beta = dtype((rho_new / rho) * (alpha / omega)) if omega != 0 else dtype(0)
        p = r + beta * (p - omega * v)
        rho = rho_new
    return x, maxiter, MAXITER, hist


def richardson(A, b, omega, iters, x0=None, dtype=np.float64):
    """``iters`` steps of x <- x + omega (b - A x) from x0 (0 by default)."""
    A = _as(A, dtype)
    b = np.asarray(b, dtype=dtype)
    om = dtype(omega)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=dtype).copy()
    for _ in range(iters):
        x = x + om * (b - A @ x)
    return x


def richardson_closed_form(A_dense, b, omega, iters):
    """x_k = sum_{j<k} (I - omega A)^j omega b, evaluated with dense matrix powers."""
    A = np.asarray(A_dense, dtype=np.float64)
    n = A.shape[0]
    G = np.eye(n) - omega * A
    acc = np.zeros(n)
    term = omega * np.asarray(b, dtype=np.float64)
    for _ in range(iters):
        acc += term
        term = G @ term
    return acc


def _scipy_solve(fn, A, b, tol, maxiter, **kw):
    it = [0]

    def cb(_):
        it[0] += 1

    x, info = fn(A, b, rtol=tol, atol=0.0, maxiter=maxiter, callback=cb, **kw)
    return x, it[0], info


def scipy_cg(A, b, tol=1e-6, maxiter=10000, jacobi=False):
    """scipy.sparse.linalg.cg (the CPU baseline), optionally preconditioned with
    M = diag(A)^-1. Returns (x, iterations, info)."""
    from scipy.sparse.linalg import LinearOperator
    from scipy.sparse.linalg import cg as _cg

    kw = {}
    if jacobi:
        dinv = 1.0 / A.diagonal()
        kw["M"] = LinearOperator(A.shape, matvec=lambda v: dinv * v, dtype=dinv.dtype)
    return _scipy_solve(_cg, A, b, tol, maxiter, **kw)


def scipy_bicgstab(A, b, tol=1e-6, maxiter=10000):
    """scipy.sparse.linalg.bicgstab (the CPU baseline). Returns (x, iterations, info)."""
    from scipy.sparse.linalg import bicgstab as _bicg

    return _scipy_solve(_bicg, A, b, tol, maxiter)
