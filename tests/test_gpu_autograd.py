"""Differentiable sparse solve against dense FP64 autograd (torch.linalg.solve on the CPU).

Bars: with float64 right-hand sides (FP64 refinement to 1e-12) x and dL/db agree
with the dense FP64 reference to 1e-8 relative (2-norm) and dL/dvalues (float32,
the dtype of the values) to 1e-6; with float32 (FP32 Krylov to 1e-6) all three
to 2e-3 relative.
"""
import numpy as np
import pytest

from gpu_util import say
from mpsparse_amd.problems import block_sparse, diffusion_2d, unpack_keys
from oracle import coo_csr as oc

pytestmark = pytest.mark.gpu


def _dense_ref(rows, cols, vals, b, w):
    import torch

    v = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    bb = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    n = b.size
    A = torch.zeros((n, n), dtype=torch.float64).index_put((torch.tensor(rows), torch.tensor(cols)), v,
                                                          accumulate=True)
    x = torch.linalg.solve(A, bb)
    (x * torch.tensor(w)).sum().backward()
    return x.detach().numpy(), v.grad.numpy(), bb.grad.numpy()


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


CASES = {
    "spd_diffusion_12": ("cg", lambda: diffusion_2d(12, sigma=1.0, seed=1)),
    "nonsym_block_8": ("bicgstab", lambda: block_sparse(8, 16, 3, seed=2)),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_sparse_solve_gradients(name, dtype):
    import torch

    from mpsparse_amd.autograd import sparse_solve

    method, make = CASES[name]
    M = make()
    if hasattr(M, "row_ptr"):
        rows = np.repeat(np.arange(M.num_rows), np.diff(M.row_ptr.astype(np.int64)))
        cols, vals, n = M.col_ind.astype(np.int64), M.vals, M.num_rows
    else:
        r, c = unpack_keys(M.keys)
        rows, cols, vals, n = r.astype(np.int64), c.astype(np.int64), M.vals, M.num_rows
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    w = rng.standard_normal(n)
    x_ref, gv_ref, gb_ref = _dense_ref(rows, cols, vals.astype(np.float64), b, w)
    td = torch.float64 if dtype == "float64" else torch.float32
    v = torch.tensor(vals, dtype=torch.float32, device="cuda", requires_grad=True)
    bt = torch.tensor(b, dtype=td, device="cuda", requires_grad=True)
    say("sparse_solve", name, dtype)
    x = sparse_solve(torch.tensor(rows, device="cuda"), torch.tensor(cols, device="cuda"), v, bt, method=method,
                     tol=1e-12 if dtype == "float64" else 1e-6)
    (x * torch.tensor(w, dtype=td, device="cuda")).sum().backward()
    bar = 1e-8 if dtype == "float64" else 2e-3
    assert _rel(x.detach().cpu().numpy(), x_ref) <= bar
    assert _rel(bt.grad.cpu().numpy(), gb_ref) <= bar
    assert _rel(v.grad.cpu().numpy(), gv_ref) <= max(bar, 1e-6)


def test_sparse_solve_rejects_bad_input():
    import torch

    from mpsparse_amd.autograd import sparse_solve

    z = torch.zeros(4, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError):
        sparse_solve(z, z, torch.ones(4, device="cuda"), torch.ones(4, device="cuda"), method="lu")


def test_sparse_solve_float64_values():
    """float64 values that FP32 holds exactly: x and every gradient (dL/dvalues
    in float64) agree with dense FP64 autograd to 1e-8."""
    import torch

    from mpsparse_amd.autograd import sparse_solve

    M = diffusion_2d(12, sigma=1.0, seed=1)
    rows = np.repeat(np.arange(M.num_rows), np.diff(M.row_ptr.astype(np.int64)))
    cols, n = M.col_ind.astype(np.int64), M.num_rows
    vals64 = M.vals.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(7)
    b, w = rng.standard_normal(n), rng.standard_normal(n)
    x_ref, gv_ref, gb_ref = _dense_ref(rows, cols, vals64, b, w)
    v = torch.tensor(vals64, dtype=torch.float64, device="cuda", requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device="cuda", requires_grad=True)
    say("sparse_solve float64 values")
    x = sparse_solve(torch.tensor(rows, device="cuda"), torch.tensor(cols, device="cuda"), v, bt, tol=1e-12)
    (x * torch.tensor(w, dtype=torch.float64, device="cuda")).sum().backward()
    assert v.grad.dtype == torch.float64
    assert _rel(x.detach().cpu().numpy(), x_ref) <= 1e-8
    assert _rel(bt.grad.cpu().numpy(), gb_ref) <= 1e-8
    assert _rel(v.grad.cpu().numpy(), gv_ref) <= 1e-8
