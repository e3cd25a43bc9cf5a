"""Direct sparse solve (banded Householder QR, FP64) against the oracle.

Bars: ||x - x_oracle|| < 1e-8 on the reference benchmark's 1D Laplacian (the
reference's own PASS criterion against scipy's LU, sparse_bridge/benchmark.py:
82-83); relative error <= 1e-10 against the dense FP64 solution elsewhere
(well-conditioned systems, backward-stable QR); least-squares optimality
|A^T (A x - b)| <= 1e-9 |A^T||b| for tall systems; gradients against dense
FP64 autograd to 1e-9 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gpu_util import say
from mpsparse_amd.problems import laplacian_1d_csc
from oracle import direct as D

pytestmark = pytest.mark.gpu


def _solver(cs, ri, v, m, n, device="cuda"):
    import torch

    from mpsparse_amd.sparse_qr import SparseQRSolver

    t = lambda a, dt: torch.tensor(np.asarray(a), dtype=dt, device=device)  # noqa: E731
    return SparseQRSolver(t(cs, torch.int64), t(ri, torch.int64), t(v, torch.float64), m, n)


def _csc(A):
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64)


@pytest.mark.parametrize("n", [1000, 4000])
def test_laplacian_1d_against_splu(n):
    import torch

    cs, ri, v = laplacian_1d_csc(n)
    A = sp.csc_matrix((v, ri, cs), shape=(n, n))
    lu = spla.splu(A)
    rng = np.random.default_rng(n)
    say("direct 1d laplacian", n)
    S = _solver(cs, ri, v, n, n)
    assert S.bandwidth == (1, 1)
    B = rng.random((n, 128))
    X = S.solve(torch.tensor(B, device="cuda")).cpu().numpy()
    assert np.linalg.norm(X - lu.solve(B)) < 1e-8
    b = rng.random(n)
    x = S.solve(torch.tensor(b)).numpy()  # CPU in, CPU out
    assert np.linalg.norm(x - lu.solve(b)) < 1e-8


def test_shuffled_nonsymmetric_band_is_recovered():
    import torch

    rng = np.random.default_rng(3)
    n, kl, ku = 600, 3, 5
    diags = [rng.standard_normal(n - abs(o)) for o in range(-kl, ku + 1)]
    A = sp.diags(diags, list(range(-kl, ku + 1)), shape=(n, n)).tolil()
    A.setdiag(A.diagonal() + 8.0)
    p = rng.permutation(n)
    A = sp.csc_matrix(A)[p][:, p]
    cs, ri, v = _csc(A)
    say("direct shuffled band")
    S = _solver(cs, ri, v, n, n)
    lo, up = S.bandwidth
    assert lo + up + 1 <= 2 * (kl + ku) + 1  # the ordering found a narrow band again
    B = rng.standard_normal((n, 3))
    X = S.solve(torch.tensor(B, device="cuda")).cpu().numpy()
    Xo = D.least_squares(cs, ri, v, n, n, B)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-10


@pytest.mark.parametrize("kl,ku", [(0, 0), (1, 2), (2, 2), (3, 4), (3, 5), (4, 3)])
def test_band_shapes(kl, ku):
    """Register-window kernels (kl + 1 <= 4, kl + ku + 1 <= 8), the LDS-window
    kernels, and mixtures of both, on sizes that are not multiples of the
    chunk lengths and 70 right-hand sides (two solve workgroups)."""
    import torch

    rng = np.random.default_rng(100 + 10 * kl + ku)
    n = 257
    offs = list(range(-kl, ku + 1))
    A = sp.diags([rng.standard_normal(n - abs(o)) for o in offs], offs, shape=(n, n)).tolil()
    A.setdiag(A.diagonal() + 6.0)
    cs, ri, v = _csc(A)
    say("direct band", kl, ku)
    S = _solver(cs, ri, v, n, n)
    lo, up = S.bandwidth
    assert lo + up <= kl + ku
    B = rng.standard_normal((n, 70))
    X = S.solve(torch.tensor(B, device="cuda")).cpu().numpy()
    Xo = D.least_squares(cs, ri, v, n, n, B)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-12


def test_laplacian_2d_grid():
    import torch

    side = 20
    T = sp.diags([-np.ones(side - 1), 4 * np.ones(side), -np.ones(side - 1)], [-1, 0, 1])
    A = (sp.kron(sp.eye(side), T) + sp.kron(sp.diags([-np.ones(side - 1), -np.ones(side - 1)], [-1, 1]),
                                               sp.eye(side))).tocsc()
    n = side * side
    cs, ri, v = _csc(A)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    say("direct 2d grid", side)
    x = _solver(cs, ri, v, n, n).solve(torch.tensor(b, device="cuda")).cpu().numpy()
    xo = D.least_squares(cs, ri, v, n, n, b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10


def test_tall_least_squares():
    import torch

    from mpsparse_amd.sparse_qr import SparseQRSolver

    # tall systems fit the banded solver when m - n plus the column spread stays
    # within the window (the reflector of column j spans rows j .. j + kl)
    rng = np.random.default_rng(5)
    m, n = 330, 300
    rows, cols, vals = [], [], []
    for j in range(n):  # column j lives on rows ~1.1 j .. 1.1 j + 3
        for r in range(int(1.1 * j), min(m, int(1.1 * j) + 4)):
            rows.append(r)
            cols.append(j)
            vals.append(rng.standard_normal() + (2.0 if r == int(1.1 * j) else 0.0))
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    cs, ri, v = _csc(A)
    b = rng.standard_normal(m)
    say("direct tall least squares")
    x = _solver(cs, ri, v, m, n).solve(torch.tensor(b, device="cuda")).cpu().numpy()
    Ad = A.toarray()
    assert np.abs(Ad.T @ (Ad @ x - b)).max() <= 1e-9 * np.abs(Ad.T).sum(axis=1).max() * np.abs(b).max()
    xo = D.least_squares(cs, ri, v, m, n, b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-9
    # much taller than wide: the band needs kl >= m - n, beyond the window
    T = sp.random(450, 300, density=0.01, random_state=11, format="csc") + sp.eye(450, 300, format="csc")
    tcs, tri, tv = _csc(T)
    with pytest.raises(RuntimeError, match="bandwidth"):
        SparseQRSolver(torch.tensor(tcs), torch.tensor(tri), torch.tensor(tv), 450, 300)


def test_duplicates_added_and_layouts():
    import torch

    n = 50
    cs, ri, v = laplacian_1d_csc(n)
    # split every diagonal entry into two duplicates
    A = sp.csc_matrix((v, ri, cs), shape=(n, n)).tocoo()
    d = A.row == A.col
    rows = np.concatenate([A.row, A.row[d]])
    cols = np.concatenate([A.col, A.col[d]])
    vals = np.concatenate([np.where(d, A.data / 2, A.data), A.data[d] / 2])
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    cs2 = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=n))]).astype(np.int64)
    say("direct duplicates")
    S = _solver(cs2, rows, vals, n, n, device="cpu")
    rng = np.random.default_rng(6)
    B = torch.tensor(rng.standard_normal((7, n))).t()  # a non-contiguous 2-D right-hand side
    X = S.solve(B)
    assert X.shape == (n, 7) and X.device.type == "cpu"
    Xo = D.least_squares(cs, ri, v, n, n, B.numpy())
    assert np.linalg.norm(X.numpy() - Xo) / np.linalg.norm(Xo) <= 1e-12


def test_singular_and_unsupported_raise():
    n = 40
    cs, ri, v = laplacian_1d_csc(n)
    v = v.copy()
    v[cs[5]:cs[6]] = 0.0  # column 5 all zero
    say("direct singular")
    with pytest.raises(RuntimeError, match="rank-deficient"):
        _solver(cs, ri, v, n, n)
    A = sp.random(300, 300, density=0.3, random_state=7, format="csc") + sp.eye(300, format="csc")
    with pytest.raises(RuntimeError, match="bandwidth"):
        _solver(*_csc(A), 300, 300)
    with pytest.raises(RuntimeError, match="num_rows >= num_cols"):
        _solver(*_csc(sp.random(10, 20, density=0.5, random_state=8, format="csc")), 10, 20)


def test_sparse_solve_gradients_match_dense():
    """SparseSolve.apply(solver, A_sparse, b) (sparse_bridge/SPQR_layer.py:12-58)
    against torch.linalg.solve + autograd in FP64 on the CPU
    (sparse_bridge/benchmark_gradients.py:55-78)."""
    import torch

    from mpsparse_amd.sparse_qr import SparseSolve, solver_from_sparse

    n = 200
    cs, ri, v = laplacian_1d_csc(n)
    rng = np.random.default_rng(9)
    v = v + 0.05 * rng.standard_normal(v.size)  # nonsymmetric values on the same pattern
    A = sp.csc_matrix((v, ri, cs), shape=(n, n)).tocoo()
    idx = torch.tensor(np.vstack([A.row, A.col]), dtype=torch.int64)
    A_sparse = torch.sparse_coo_tensor(idx, torch.tensor(A.data), (n, n)).coalesce().requires_grad_(True)
    b = torch.tensor(rng.standard_normal(n), requires_grad=True)
    target = torch.tensor(rng.standard_normal(n))
    say("direct autograd")
    x = SparseSolve.apply(solver_from_sparse(A_sparse.detach()), A_sparse, b)
    torch.nn.functional.mse_loss(x, target).backward()
    Ad = A_sparse.detach().to_dense().clone().requires_grad_(True)
    bd = b.detach().clone().requires_grad_(True)
    xd = torch.linalg.solve(Ad, bd)
    torch.nn.functional.mse_loss(xd, target).backward()
    assert torch.linalg.vector_norm(x.detach() - xd.detach()) / torch.linalg.vector_norm(xd) <= 1e-10
    assert torch.linalg.vector_norm(b.grad - bd.grad) / torch.linalg.vector_norm(bd.grad) <= 1e-9
    ga = A_sparse.grad.coalesce()
    gd = Ad.grad[ga.indices()[0], ga.indices()[1]]
    assert torch.linalg.vector_norm(ga.values() - gd) / torch.linalg.vector_norm(gd) <= 1e-9


def test_get_grad_a_helper():
    import torch

    n = 30
    cs, ri, v = laplacian_1d_csc(n)
    S = _solver(cs, ri, v, n, n)
    rng = np.random.default_rng(10)
    gb, x = rng.standard_normal(n), rng.standard_normal(n)
    say("direct grad helper")
    g = S.get_grad_a_helper(torch.tensor(cs), torch.tensor(ri), torch.tensor(gb), torch.tensor(x), n, n).numpy()
    cols = np.repeat(np.arange(n), np.diff(cs))
    assert np.array_equal(g, -gb[ri] * x[cols])
