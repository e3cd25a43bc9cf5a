"""GPU BCSR: conversion (COO -> BCSR on the GPU) against the oracle's CSR -> BSR,
SpMV against the FP64 block product, and the solvers.

Conversion: brow_ptr / bcol bit-exact; blocks bit-exact without duplicate keys
(with duplicates: the GPU sums in FP32 in input order, the oracle in FP64 --
compared to 1e-6 relative). SpMV bound: 2 * gamma_k |A||x| as for CSR.
"""
import numpy as np
import pytest

from gpu_util import check_fp32_solution, say, t_f32, t_keys
from mpsparse_amd.problems import block_sparse, diffusion_2d, laplacian_2d, pack_keys
from oracle import coo_csr as oc
from oracle import krylov as K
from oracle import spmv as S

pytestmark = pytest.mark.gpu


def _from_coo(keys, vals, n, m, B):
    import torch

    from mpsparse_amd import bcsr

    return bcsr.bcsr_tensor(t_keys(keys), t_f32(vals), torch.empty(0, dtype=torch.int32),
                            torch.empty(0, dtype=torch.int32), n, m, B)


def _oracle_bsr(keys, vals, n, m, B):
    rp, ci, v = oc.coo_to_csr(keys, vals, n, m)
    return (rp, ci, v), S.csr_to_bsr(rp, ci, v, n, m, B)


@pytest.mark.parametrize("B", [4, 8, 16])
@pytest.mark.parametrize("case", ["laplacian", "random_partial", "duplicates"])
def test_coo_to_bcsr_matches_oracle(B, case):
    rng = np.random.default_rng(B * 10 + len(case))
    if case == "laplacian":
        p = laplacian_2d(48)
        keys, vals, n, m = p.keys, p.vals, p.num_rows, p.num_cols
    elif case == "random_partial":
        n, m = 1000, 1003
        rows = rng.integers(0, n, 30000)
        cols = rng.integers(0, m, 30000)
        keys = np.unique(pack_keys(rows, cols))
        rng.shuffle(keys)
        vals = rng.standard_normal(keys.size).astype(np.float32)
    else:
        n, m = 300, 300
        keys = pack_keys(rng.integers(0, n, 20000), rng.integers(0, m, 20000))
        vals = rng.standard_normal(keys.size).astype(np.float32)
    say("coo->bcsr", case, B)
    A = _from_coo(keys, vals, n, m, B)
    (rp, ci, v), (brp, bci, blocks, nbr, nbc) = _oracle_bsr(keys, vals, n, m, B)
    g_rp, g_bc, g_bl = (t.cpu().numpy() for t in A.export())
    assert A.num_blocks == bci.size
    assert np.array_equal(g_rp.view(np.uint32), brp)
    assert np.array_equal(g_bc.view(np.uint32), bci)
    if case == "duplicates":
        assert np.allclose(g_bl, blocks, rtol=1e-6, atol=1e-6)
    else:
        assert np.array_equal(g_bl, blocks)


@pytest.mark.parametrize("B", [4, 8, 16])
def test_bcsr_spmv(B):
    import torch

    q = block_sparse(512, 16, 8, seed=3)
    say("bcsr spmv", B)
    A = _from_coo(q.keys, q.vals, q.num_rows, q.num_cols, B)
    (rp, ci, v), _ = _oracle_bsr(q.keys, q.vals, q.num_rows, q.num_cols, B)
    x = np.random.default_rng(B).standard_normal(q.num_cols).astype(np.float32)
    y = torch.full((q.num_rows,), np.nan, device="cuda")
    A.mv(t_f32(x), y)
    y64 = S.csr_spmv(rp, ci, v, x)
    bound = 2 * S.spmv_error_bound(rp, ci, v, x) + 1e-30
    assert np.all(np.abs(y.cpu().numpy() - y64) <= bound)


def test_reference_calling_convention():
    """Host BSR arrays (scipy tobsr layout), b padded, x logical -- bcsr_benchmark.py:70-116."""
    import torch

    from mpsparse_amd import bcsr

    p = laplacian_2d(31)  # n = 961: the last block row is partial
    n = p.num_rows
    (rp, ci, v), (brp, bci, blocks, nbr, nbc) = _oracle_bsr(p.keys, p.vals, n, n, 4)
    A = bcsr.bcsr_tensor(torch.empty(0), torch.from_numpy(blocks.reshape(-1).copy()),
                         torch.from_numpy(brp.view(np.int32).copy()), torch.from_numpy(bci.view(np.int32).copy()),
                         n, n, 4)
    b_pad = torch.zeros(nbr * 4)
    b_pad[:n] = torch.from_numpy(p.b)
    x = torch.zeros(n)
    say("bcsr reference convention")
    info = A.bicgstab(b_pad, x)
    Asp = S_csr(rp, ci, v, n)
    assert info["status"] == "converged" and K.true_relres(Asp, x.numpy(), p.b) <= 1.05e-6, info


def S_csr(rp, ci, v, n):
    import scipy.sparse as sp

    return sp.csr_matrix((v.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))


def test_bcsr16_bicgstab_block_sparse():
    q = block_sparse(1024, 16, 8, seed=6)
    say("bcsr16 bicgstab")
    A = _from_coo(q.keys, q.vals, q.num_rows, q.num_cols, 16)
    assert A.num_blocks == 1024 * 9
    (rp, ci, v), _ = _oracle_bsr(q.keys, q.vals, q.num_rows, q.num_cols, 16)
    import torch

    x = torch.zeros(q.num_rows, device="cuda")
    info = A.bicgstab(t_f32(q.b), x)
    Asp = S_csr(rp, ci, v, q.num_rows)
    assert info["status"] == "converged" and K.true_relres(Asp, x.cpu().numpy(), q.b) <= 1.05e-6, info
    _, it64, _, _ = K.bicgstab(Asp, q.b.astype(np.float64), tol=1e-6)
    assert abs(info["iterations"] - it64) <= 0.15 * it64 + 2


def test_bcsr_jacobi_cg_and_richardson():
    import torch

    M = diffusion_2d(50, sigma=1.0, seed=7)  # n = 2500
    rows = np.repeat(np.arange(M.num_rows), np.diff(M.row_ptr.astype(np.int64)))
    keys = pack_keys(rows, M.col_ind)
    say("bcsr jacobi cg")
    A = _from_coo(keys, M.vals, M.num_rows, M.num_cols, 8)
    b = np.random.default_rng(1).standard_normal(M.num_rows).astype(np.float32)
    x = torch.zeros(M.num_rows, device="cuda")
    info = A.cgd(t_f32(b), x, precond="jacobi")
    Asp = M.to_scipy()
    check_fp32_solution(info, Asp, x.cpu().numpy(), b, 1e-6)
    omega = 1.0 / float(abs(Asp).sum(axis=1).max())
    x = torch.zeros(M.num_rows, device="cuda")
    A.richardson(t_f32(b), x, omega=omega, iters=300)
    x64 = K.richardson(Asp, b.astype(np.float64), omega, 300)
    assert np.max(np.abs(x.cpu().numpy() - x64)) <= 1e-4 * np.max(np.abs(x64))
