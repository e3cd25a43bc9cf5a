"""Oracle of the direct sparse solve, pinned on the CPU: the dense FP64
restatement against scipy's sparse LU (the reference benchmark's own check,
sparse_bridge/benchmark.py:62-86), manufactured solutions, and the benchmark
matrix against its formula."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mpsparse_amd.problems import laplacian_1d_csc
from oracle import direct as D


def test_benchmark_matrix_matches_reference_formula():
    for n in (1, 2, 7, 50):
        cs, ri, v = laplacian_1d_csc(n)
        ocs, ori, ov = D.laplacian_1d_shifted_csc(n)
        assert np.array_equal(cs, ocs) and np.array_equal(ri, ori) and np.array_equal(v, ov)
        A = D.dense_from_csc(cs, ri, v, n, n)
        assert np.allclose(np.diag(A), 2.1) and np.allclose(np.diag(A, 1), -1) and np.allclose(np.diag(A, -1), -1)


def test_square_matches_splu_and_manufactured_solution():
    n = 300
    cs, ri, v = laplacian_1d_csc(n)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((n, 5))
    A = sp.csc_matrix((v, ri, cs), shape=(n, n))
    B = A @ X
    x = D.least_squares(cs, ri, v, n, n, B)
    assert np.linalg.norm(x - X) / np.linalg.norm(X) < 1e-12
    assert np.linalg.norm(x - spla.splu(A).solve(B)) < 1e-8


def test_tall_least_squares_normal_equations():
    rng = np.random.default_rng(1)
    m, n = 40, 25
    A = sp.random(m, n, density=0.3, random_state=2, format="csc") + sp.eye(m, n, format="csc")
    A = A.tocsc()
    b = rng.standard_normal(m)
    x = D.least_squares(A.indptr, A.indices, A.data, m, n, b)
    Ad = A.toarray()
    # optimality: the residual is orthogonal to the columns of A
    assert np.abs(Ad.T @ (Ad @ x - b)).max() < 1e-10
    # a plausible mistake (dropping the least-squares step, solving the top square block) fails
    x_bad = np.linalg.solve(Ad[:n], b[:n])
    assert np.abs(Ad.T @ (Ad @ x_bad - b)).max() > 1e-3


def test_bandwidths():
    cs, ri, v = laplacian_1d_csc(10)
    assert D.bandwidths(cs, ri, 10) == (1, 1)
    p = np.arange(10)[::-1].copy()
    assert D.bandwidths(cs, ri, 10, p, p) == (1, 1)
