"""Pins the SpMV and Krylov oracles with closed forms and independent routines (CPU only).

* SpMV: dense numpy products; the 2D Laplacian's eigenvectors
  (sin-products, eigenvalue 4 - 2cos(p pi/65) - 2cos(q pi/65)); the scipy BSR
  conversion; the C sequential FP32 sum within the gamma_k error bound.
* CG: exact one-step solve for a diagonal matrix; termination in <= k steps for
  k distinct eigenvalues; the C1 iteration bound from kappa; scipy / spsolve.
* BiCGSTAB: spsolve and scipy iteration counts on a nonsymmetric system.
* Richardson: the closed form sum_j (I - omega A)^j omega b.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mpsparse_amd import problems as P
from oracle import coo_csr as oc
from oracle import krylov as K
from oracle import spmv as S


def _random_csr(rng, n, m, nnz):
    rows = rng.integers(0, n, nnz)
    cols = rng.integers(0, m, nnz)
    vals = rng.standard_normal(nnz).astype(np.float32)
    return oc.coo_to_csr(P.pack_keys(rows, cols), vals, n, m)


def test_csr_spmv_matches_dense():
    rng = np.random.default_rng(1)
    n, m = 57, 41
    rp, ci, v = _random_csr(rng, n, m, 400)  # duplicates and empty rows included
    dense = np.zeros((n, m))
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    np.add.at(dense, (rows, ci.astype(np.int64)), v.astype(np.float64))
    x = rng.standard_normal(m)
    assert np.allclose(S.csr_spmv(rp, ci, v, x), dense @ x, rtol=1e-13, atol=1e-13)


def test_laplacian_eigenvectors():
    A = P.laplacian_2d_csr(64)
    g = 64
    i = np.arange(1, g + 1)
    for p_, q_ in [(1, 1), (3, 7), (64, 64), (10, 55)]:
        # row index i*64 + j holds grid point (i, j)
        x = np.outer(np.sin(i * p_ * np.pi / (g + 1)), np.sin(i * q_ * np.pi / (g + 1))).ravel()
        lam = 4 - 2 * np.cos(p_ * np.pi / (g + 1)) - 2 * np.cos(q_ * np.pi / (g + 1))
        assert np.allclose(S.csr_spmv(A.row_ptr, A.col_ind, A.vals, x), lam * x, atol=1e-12)


def test_bsr_conversion_matches_scipy():
    rng = np.random.default_rng(2)
    n, m, B = 50, 45, 4
    rp, ci, v = _random_csr(rng, n, m, 300)
    brp, bci, blocks, nbr, nbc = S.csr_to_bsr(rp, ci, v, n, m, B)
    A = sp.csr_matrix((v.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(n, m))
    indptr = np.concatenate([A.indptr, np.full(nbr * B - n, A.indptr[-1])])
    Ap = sp.csr_matrix((A.data, A.indices, indptr), shape=(nbr * B, nbc * B))
    bsr = Ap.tobsr(blocksize=(B, B))
    bsr.sum_duplicates()
    bsr.sort_indices()
    assert np.array_equal(brp.astype(np.int64), bsr.indptr)
    assert np.array_equal(bci.astype(np.int64), bsr.indices)
    assert np.allclose(blocks, bsr.data, atol=1e-6)
    x = rng.standard_normal(nbc * B)
    assert np.allclose(S.bsr_spmv(brp, bci, blocks, x), Ap @ x, atol=1e-5)


def test_fp32_sequential_spmv_within_bound():
    rng = np.random.default_rng(3)
    rp, ci, v = _random_csr(rng, 300, 300, 6000)
    x = rng.standard_normal(300).astype(np.float32)
    y32 = oc.spmv_f32_seq(rp, ci, v, x)
    y64 = S.csr_spmv(rp, ci, v, x)
    bound = S.spmv_error_bound(rp, ci, v, x)
    assert np.all(np.abs(y32 - y64) <= bound + 1e-30)


def test_cg_diagonal_one_step():
    d = np.array([1.0, 2.0, 5.0, 0.5])
    A = sp.diags(d).tocsr()
    b = np.array([1.0, -2.0, 3.0, 4.0])
    x, it, st, _ = K.cg(A, b, tol=1e-12)
    # diagonal with distinct entries is not one step in general; use a scalar multiple of I
    A1 = sp.diags(np.full(4, 3.0)).tocsr()
    x1, it1, st1, _ = K.cg(A1, b, tol=1e-14)
    assert it1 == 1 and st1 == K.CONVERGED and np.allclose(x1, b / 3.0, rtol=0, atol=1e-15)
    assert st == K.CONVERGED and it <= 4 and np.allclose(x, b / d)


def test_cg_terminates_in_distinct_eigenvalue_count():
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    lam = np.repeat([1.0, 4.0, 9.0], 10)
    A = sp.csr_matrix(Q @ np.diag(lam) @ Q.T)
    b = rng.standard_normal(30)
    x, it, st, _ = K.cg(A, b, tol=1e-10)
    assert st == K.CONVERGED and it <= 3


def test_cg_laplacian_bound_and_scipy():
    p = P.laplacian_2d(64)
    A = P.laplacian_2d_csr(64).to_scipy()
    lam_min = 4 - 4 * np.cos(np.pi / 65)
    lam_max = 4 + 4 * np.cos(np.pi / 65)
    kappa = lam_max / lam_min
    assert abs(kappa - 1711.7) < 0.5
    sk = np.sqrt(kappa)
    bound = np.log(2 * sk / 1e-6) / np.log((sk + 1) / (sk - 1))
    x, it, st, _ = K.cg(A, p.b.astype(np.float64), tol=1e-6)
    assert st == K.CONVERGED and it <= bound
    assert K.true_relres(A, x, p.b) <= 1e-6
    xs = spla.spsolve(A.tocsc(), p.b.astype(np.float64))
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= kappa * 1e-6
    _, it_sp, info = K.scipy_cg(A, p.b.astype(np.float64), tol=1e-6)
    assert info == 0 and abs(it_sp - it) <= 2


def test_cg_fp32_matches_fp64_iterations():
    p = P.laplacian_2d(64)
    A = P.laplacian_2d_csr(64).to_scipy()
    _, it64, _, _ = K.cg(A, p.b, tol=1e-6)
    x32, it32, st, _ = K.cg(A, p.b, tol=1e-6, dtype=np.float32)
    assert st == K.CONVERGED and abs(it32 - it64) <= 0.1 * it64 + 2
    # plain FP32 recurrences drift: the true residual of the FP32 iterate stalls
    # near 4e-6 on this system although the recurrence reports < 1e-6 (the reason
    # the GPU solvers re-check convergence with a compensated true residual)
    assert 1e-6 < K.true_relres(A, x32, p.b) <= 1e-5


def test_bicgstab_nonsymmetric():
    q = P.block_sparse(64, 16, 8, seed=5)
    rp, ci, v = oc.coo_to_csr(q.keys, q.vals, q.num_rows, q.num_cols)
    A = sp.csr_matrix((v.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(q.num_rows,) * 2)
    b = q.b.astype(np.float64)
    x, it, st, _ = K.bicgstab(A, b, tol=1e-6)
    assert st == K.CONVERGED
    assert K.true_relres(A, x, b) <= 1e-6
    xs = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= 1e-5
    _, it_sp, info = K.scipy_bicgstab(A, b, tol=1e-6)
    assert info == 0 and abs(it_sp - it) <= 2


def test_pcg_jacobi_on_diffusion():
    A = P.diffusion_2d(24, sigma=1.5, seed=2).to_scipy()
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    dinv = 1.0 / A.diagonal()
    x, it, st, _ = K.cg(A, b, tol=1e-8, M_apply=lambda r: dinv * r)
    _, it_plain, _, _ = K.cg(A, b, tol=1e-8)
    assert st == K.CONVERGED and K.true_relres(A, x, b) <= 1e-8
    assert it < it_plain


def test_richardson_closed_form():
    rng = np.random.default_rng(6)
    M = rng.standard_normal((8, 8))
    Ad = M @ M.T + 8 * np.eye(8)
    A = sp.csr_matrix(Ad)
    b = rng.standard_normal(8)
    omega = 1.0 / np.linalg.eigvalsh(Ad).max()
    for iters in (0, 1, 5, 40):
        x = K.richardson(A, b, omega, iters)
        assert np.allclose(x, K.richardson_closed_form(Ad, b, omega, iters), rtol=1e-12, atol=1e-12)


def test_round16_rowsum_pins():
    """The 16-bit value copy keeps A's row sums, rounds off-diagonals plainly and,
    on a diffusion M-matrix, changes the energy by at most the unit roundoff."""
    from mpsparse_amd.problems import diffusion_3d
    from oracle import spmv as SP

    M = diffusion_3d(6, 5, 4, sigma=2.0, seed=1)
    rp, ci, v = M.row_ptr.astype(np.int64), M.col_ind.astype(np.int64), M.vals
    n = M.num_rows
    rows = np.repeat(np.arange(n), np.diff(rp))
    A = sp.csr_matrix((v.astype(np.float64), ci, rp), shape=(n, n))
    rng = np.random.default_rng(0)
    for fmt, u in (("bf16", 2.0 ** -8), ("f16", 2.0 ** -11)):
        v16, dlo = SP.round16_rowsum(rp, ci, v, fmt)
        A16 = sp.csr_matrix((v16, ci, rp), shape=(n, n)) + sp.diags(dlo)
        rs, rs16 = np.asarray(A.sum(1)).ravel(), np.asarray(A16.sum(1)).ravel()
        assert np.all(np.abs(rs16 - rs) <= 1e-6 * np.asarray(abs(A).sum(1)).ravel())
        off = ci != rows
        plain = SP.round_bf16(v) if fmt == "bf16" else v.astype(np.float16).astype(np.float32)
        assert np.array_equal(v16[off], plain[off].astype(np.float64))
        for _ in range(5):
            x = rng.standard_normal(n)
            e, e16 = x @ (A @ x), x @ (A16 @ x)
            assert abs(e16 - e) <= u * e, (fmt, e, e16)
        # the SpMV the kernels compute
        y = SP.csr_spmv(rp, ci, v16, x) + dlo * x
        assert np.allclose(y, A16 @ x, rtol=0, atol=1e-12 * np.abs(A16) @ np.abs(x))
    # duplicate diagonal entries and a row without a diagonal
    rp2 = np.array([0, 3, 4])
    ci2 = np.array([0, 0, 1, 0])
    v2 = np.array([2.0, 1.0, -0.3333333, 5.0], dtype=np.float32)
    v16, dlo = SP.round16_rowsum(rp2, ci2, v2, "bf16")
    assert v16[1] == 0.0 and v16[0] == SP.round_bf16(np.float32(3.0 + (-0.3333333 - SP.round_bf16(-0.3333333))))
    assert abs(v16[0] + dlo[0] + v16[2] - (3.0 - 0.3333333)) < 1e-6
    assert v16[3] == SP.round_bf16(5.0) and dlo[1] == (5.0 - SP.round_bf16(5.0))
