"""Diagonal storage of stencil-like matrices (format 'dia', chosen automatically).

Bars: SpMV within 2 * gamma_k (|A||x|) of the FP64 oracle product (the same bound
as the CSR kernel; only the summation order differs), 16-bit values against the
oracle's row-sum-preserving rounding (oracle.spmv.round16_rowsum), solves to
the FP64 true residual 1.05 * tol with iteration counts within 5% + 2 of the CSR
path (the summation order differs).
"""
import numpy as np
import pytest

from gpu_util import build_csr, csr_from_arrays, say, t_f32
from mpsparse_amd.problems import block_sparse, diffusion_3d, laplacian_2d, pack_keys
from oracle import krylov as K
from oracle import spmv as S

pytestmark = pytest.mark.gpu


def _mv(A, x, values="f32"):
    import torch

    y = torch.full((A.num_rows,), np.nan, dtype=torch.float32, device="cuda")
    A.mv(t_f32(x), y, values=values)
    return y.cpu().numpy().astype(np.float64)


def _bound(rp, ci, v, x):
    return 2 * S.spmv_error_bound(rp, ci, v, x) + 1e-30


def test_stencil_uses_dia_and_matches_csr():
    M = diffusion_3d(20, sigma=1.5, seed=2)
    A = csr_from_arrays(M.row_ptr, M.col_ind, M.vals, M.num_rows, M.num_cols)
    x = np.random.default_rng(0).standard_normal(M.num_cols).astype(np.float32)
    say("dia spmv diffusion 20^3")
    y = _mv(A, x)
    assert A.format == "dia_symmetric"
    y64 = S.csr_spmv(M.row_ptr, M.col_ind, M.vals, x)
    assert np.all(np.abs(y - y64) <= _bound(M.row_ptr, M.col_ind, M.vals, x))
    A.set_format("csr")
    yc = _mv(A, x)
    assert A.format == "csr"
    assert np.all(np.abs(yc - y64) <= _bound(M.row_ptr, M.col_ind, M.vals, x))
    A.set_format("auto")
    assert np.array_equal(_mv(A, x), y) and A.format == "dia_symmetric"


@pytest.mark.parametrize("values", ["bf16", "f16"])
def test_dia_16bit_values(values):
    M = diffusion_3d(16, sigma=2.0, seed=4)
    A = csr_from_arrays(M.row_ptr, M.col_ind, M.vals, M.num_rows, M.num_cols)
    x = np.random.default_rng(1).standard_normal(M.num_cols).astype(np.float32)
    say("dia spmv", values)
    y = _mv(A, x, values)
    assert A.format in ("dia", "dia_symmetric")
    v16, dlo = S.round16_rowsum(M.row_ptr, M.col_ind, M.vals, values)
    y64 = S.csr_spmv(M.row_ptr, M.col_ind, v16, x) + dlo * x
    bound = _bound(M.row_ptr, M.col_ind, v16, x) + 4 * 2.0 ** -24 * np.abs(dlo * x)
    assert np.all(np.abs(y - y64) <= bound)


def test_random_pattern_stays_csr():
    q = block_sparse(64, 16, 8, seed=3)
    A, rp, ci, v = build_csr(q.keys, q.vals, q.num_rows, q.num_cols)
    x = np.random.default_rng(2).standard_normal(q.num_cols).astype(np.float32)
    say("dia probe on block-sparse")
    y = _mv(A, x)
    assert A.format == "csr"
    assert np.all(np.abs(y - S.csr_spmv(rp, ci, v, x)) <= _bound(rp, ci, v, x))


def test_duplicates_and_rectangular_bands():
    """Duplicate diagonal entries are summed; rectangular banded matrices (n < m and n > m)."""
    rng = np.random.default_rng(5)
    for n, m, offs in ((300, 340, (0, 3, 40)), (340, 300, (-40, -1, 0, 7))):
        rows, cols = [], []
        for o in offs:
            i = np.arange(n)
            keep = (i + o >= 0) & (i + o < m)
            rows.append(i[keep])
            cols.append(i[keep] + o)
        rows = np.concatenate(rows + [np.array([5, 5])])  # (5, 5) twice more: duplicates
        cols = np.concatenate(cols + [np.array([5, 5])])
        vals = rng.standard_normal(rows.size).astype(np.float32)
        keys = pack_keys(rows.astype(np.uint64), cols.astype(np.uint64))
        A, rp, ci, v = build_csr(keys, vals, n, m)
        x = rng.standard_normal(m).astype(np.float32)
        say("dia rectangular", n, m, offs)
        y = _mv(A, x)
        assert A.format == "dia"
        assert np.all(np.abs(y - S.csr_spmv(rp, ci, v, x)) <= _bound(rp, ci, v, x))
        y16 = _mv(A, x, "bf16")
        v16, dlo = S.round16_rowsum(rp, ci, v, "bf16", num_cols=m)
        k = min(n, m)
        dx = np.zeros(n)
        dx[:k] = dlo[:k] * x[:k]
        ref = S.csr_spmv(rp, ci, v16, x) + dx
        bound = _bound(rp, ci, v16, x) + 4 * 2.0 ** -24 * np.abs(dx)
        assert np.all(np.abs(y16 - ref) <= bound)


@pytest.mark.parametrize("method", ["cgd", "bicgstab"])
def test_solvers_same_on_dia_and_csr(method):
    import torch

    p = laplacian_2d(48)
    A, rp, ci, v = build_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    import scipy.sparse as sp

    Ssp = sp.csr_matrix((v.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(p.num_rows,) * 2)
    its = {}
    for fmt in ("auto", "csr"):
        A.set_format(fmt)
        x = torch.zeros(p.num_rows, device="cuda")
        say("solve", method, fmt)
        info = getattr(A, method)(t_f32(p.b), x)
        assert info["status"] == "converged", info
        assert K.true_relres(Ssp, x.cpu().numpy(), p.b) <= 1.05e-6
        its[fmt] = info["iterations"]
        assert A.format == ("dia_symmetric" if fmt == "auto" else "csr")
    assert abs(its["auto"] - its["csr"]) <= 0.05 * its["csr"] + 2, its


@pytest.mark.parametrize("values", ["f32", "bf16", "f16"])
def test_symmetric_storage_is_bit_identical(values):
    """A symmetric stencil whose value planes exceed 16 MB (216^3: 282 MB in
    FP32, 141 MB in 16 bits) reads its lower diagonals from the upper planes
    (4 value planes instead of 7 for the 7-point operator). Against the same
    matrix with one entry of its last row changed (no longer symmetric: every
    diagonal read from its own plane) the products of all other rows agree
    bit for bit, and both stay within the FP64 bound."""
    M = diffusion_3d(216, sigma=1.5, seed=7)
    n = M.num_rows
    A = csr_from_arrays(M.row_ptr, M.col_ind, M.vals, n, M.num_cols)
    v2 = M.vals.copy()
    v2[M.row_ptr[n - 1]] *= 1.5  # first entry of the last row: a lower-triangle coupling
    B = csr_from_arrays(M.row_ptr, M.col_ind, v2, n, M.num_cols)
    x = np.random.default_rng(3).standard_normal(M.num_cols).astype(np.float32)
    say("dia symmetric storage", values)
    ya, yb = _mv(A, x, values), _mv(B, x, values)
    assert A.format == "dia_symmetric" and B.format == "dia"
    assert np.array_equal(ya[:-1].view(np.uint64), yb[:-1].view(np.uint64))
    if values == "f32":
        for vv, y in ((M.vals, ya), (v2, yb)):
            y64 = S.csr_spmv(M.row_ptr, M.col_ind, vv, x)
            assert np.all(np.abs(y - y64) <= _bound(M.row_ptr, M.col_ind, vv, x))
