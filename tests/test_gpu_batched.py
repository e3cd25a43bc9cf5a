"""Batched Jacobi-PCG against the FP64 oracle, system by system.

Bars (FP32 device arithmetic, compensated iterate): with the FP64 result every
(system, rhs) pair converges to a true relative residual ||b - A x|| / ||b||
<= 1.05 * tol in FP64 (oracle), the reported per-pair relres agrees with it to
1%, and the iteration count lies between 0.85x - 3 and 1.5x + 5 of FP64
Jacobi-PCG (oracle.krylov.cg with M = D^-1) on the same system and right-hand
side (the upper slack covers true-residual restarts). With the FP32 result a
pair either converges the same way or stops as stagnated where the FP32
rounding of the exact solution is within 3x of tol, at most 4x that floor.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from gpu_util import say
from oracle import krylov as K

pytestmark = pytest.mark.gpu


def _build(num_systems, lo, hi, sigma=1.5, seed=3):
    import torch

    from mpsparse_amd import batched, spmv

    keys, vals, off, sides = batched.batched_diffusion(num_systems, lo, hi, sigma, seed)
    n = int(off[-1])
    rp = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    ci = torch.empty(keys.numel(), dtype=torch.int32, device="cuda")
    cv = torch.empty(keys.numel(), dtype=torch.float32, device="cuda")
    A = spmv.csr_tensor(keys, vals, rp, ci, cv, n, n)
    S = sp.csr_matrix((cv.cpu().numpy().astype(np.float64), ci.cpu().numpy().astype(np.int64),
                       rp.cpu().numpy().astype(np.int64)), shape=(n, n))
    return A, S, off, sides


def _check_pairs(S, off, b, x, info, tol, nrhs, its_check=True):
    import scipy.sparse.linalg as spla

    st, its, rel = (info[k].numpy() for k in ("status", "iterations", "relres"))
    fp32 = x.dtype == np.float32
    for s in range(len(off) - 1):
        lo, hi = int(off[s]), int(off[s + 1])
        Ss = S[lo:hi, lo:hi]
        d = Ss.diagonal()
        for k in range(nrhs):
            bs = b[lo:hi, k].astype(np.float64)
            if not np.any(bs):
                assert st[s, k] == 1 and its[s, k] == 0 and not np.any(x[lo:hi, k])
                continue
            r = K.true_relres(Ss, x[lo:hi, k], bs)
            assert abs(rel[s, k] - r) <= 0.01 * r + 1e-12, (x.dtype, s, k, st[s, k], its[s, k], rel[s, k], r,
                                                            st.tolist(), its.tolist())
            if st[s, k] != 1 and fp32:
                floor = K.true_relres(Ss, spla.spsolve(Ss.tocsc(), bs).astype(np.float32), bs)
                assert st[s, k] == 4 and floor > tol / 3 and r <= 4 * floor, (s, k, st[s, k], r, floor)
                continue
            assert st[s, k] == 1, (s, k, st[s, k], its[s, k], r)
            assert r <= 1.05 * tol, (s, k, r)
            if its_check:
                _, it64, _, _ = K.cg(Ss, bs, tol=tol, M_apply=lambda v, d=d: v / d)
                assert 0.85 * it64 - 3 <= its[s, k] <= 1.5 * it64 + 5, (s, k, its[s, k], it64)


@pytest.mark.parametrize("nrhs", [4, 1, 3])
def test_batched_matches_oracle(nrhs):
    import torch

    from mpsparse_amd import batched

    A, S, off, sides = _build(12, 4, 40)
    n = int(off[-1])
    B = batched.batch_tensor(A, torch.from_numpy(off))
    assert B.num_systems == 12
    b = np.random.default_rng(nrhs).standard_normal((n, nrhs)).astype(np.float32)
    b[off[2]:off[3], 0] = 0.0  # a right-hand side that is already solved
    x = torch.full((n, nrhs), np.nan, dtype=torch.float64, device="cuda")
    say("batched pcg", nrhs, "rhs, sides", sides.tolist())
    info = B.pcg(torch.from_numpy(b).cuda(), x)
    assert info["converged_pairs"] == info["total_pairs"] == 12 * nrhs, (info["status"], info["relres"])
    _check_pairs(S, off, b, x.cpu().numpy(), info, 1e-6, nrhs)
    # the FP32 result: converged, or stagnated at the FP32 rounding floor
    x32 = torch.full((n, nrhs), np.nan, dtype=torch.float32, device="cuda")
    info = B.pcg(torch.from_numpy(b).cuda(), x32)
    _check_pairs(S, off, b, x32.cpu().numpy(), info, 1e-6, nrhs, its_check=False)


def test_batched_fixed_schedules_agree():
    """Every tile size / rows-per-thread choice converges (the tuner only picks speed)."""
    import torch

    from mpsparse_amd import batched

    A, S, off, _ = _build(6, 16, 48, seed=5)
    n = int(off[-1])
    B = batched.batch_tensor(A, torch.from_numpy(off))
    b = np.random.default_rng(0).standard_normal((n, 4)).astype(np.float32)
    for T, rpt in ((256, 1), (512, 2), (2048, 1)):
        x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
        say("batched schedule", T, rpt)
        info = B.pcg(torch.from_numpy(b).cuda(), x, tile_rows=T, rows_per_thread=rpt, steps_per_graph=8)
        assert info["tile_rows"] == T and info["rows_per_thread"] == rpt
        _check_pairs(S, off, b, x.cpu().numpy(), info, 1e-6, 4, its_check=False)


def test_batched_maxiter_and_host_vectors():
    import torch

    from mpsparse_amd import batched

    A, S, off, _ = _build(5, 20, 60, seed=7)
    n = int(off[-1])
    B = batched.batch_tensor(A, torch.from_numpy(off))
    b = np.random.default_rng(2).standard_normal((n, 2)).astype(np.float32)
    x = torch.zeros((n, 2), dtype=torch.float64)
    say("batched maxiter 6, host vectors")
    info = B.pcg(torch.from_numpy(b), x, maxiter=6)
    st, its = info["status"].numpy(), info["iterations"].numpy()
    assert np.all((st == 3) & (its == 6)), (st, its)
    # the same handle solves fully afterwards (graph rebuilt for the new maxiter)
    info = B.pcg(torch.from_numpy(b), x)
    assert info["converged_pairs"] == 10
    _check_pairs(S, off, b, x.numpy(), info, 1e-6, 2, its_check=False)


def test_batched_rejects_coupled_systems():
    import torch

    from mpsparse_amd import batched
    from mpsparse_amd.problems import laplacian_2d

    from gpu_util import build_csr

    p = laplacian_2d(8)  # 64 rows; cutting at 30 splits coupled rows
    A = build_csr(p.keys, p.vals, p.num_rows, p.num_cols, export=False)[0]
    say("batched rejects coupled systems")
    with pytest.raises(RuntimeError, match="block diagonal"):
        batched.batch_tensor(A, torch.tensor([0, 30, 64]))
    with pytest.raises(RuntimeError, match="sys_off"):
        batched.batch_tensor(A, torch.tensor([0, 30, 60]))
    B = batched.batch_tensor(A, torch.tensor([0, 64]))
    assert B.num_systems == 1


def _block_diag(blocks):
    """A csr_tensor of the block-diagonal matrix of scipy SPD blocks, its scipy
    form and the system offsets."""
    import torch

    from gpu_util import csr_from_arrays

    S = sp.block_diag(blocks, format="csr").astype(np.float32)
    S.sort_indices()
    off = np.concatenate([[0], np.cumsum([b.shape[0] for b in blocks])]).astype(np.int64)
    A = csr_from_arrays(S.indptr.astype(np.uint32), S.indices.astype(np.uint32), S.data, S.shape[0], S.shape[1])
    return A, S.astype(np.float64), off, torch


def _spd_1d(n, seed, shift=0.2):
    """1D log-normal diffusion plus a diagonal shift (a moderate condition number)."""
    rng = np.random.default_rng(seed)
    k = np.exp(rng.standard_normal(n + 1))
    return sp.diags([-k[1:n], k[:n] + k[1:] + shift, -k[1:n]], [-1, 0, 1]).tocsr()


def _spd_9pt(side, seed):
    """2D 9-point operator (nine diagonals): the Kronecker sum of 1D operators plus diagonal couplings."""
    T = _spd_1d(side, seed)
    I = sp.eye(side)
    C = sp.diags([np.full(side - 1, -0.1), np.full(side - 1, -0.1)], [-1, 1])
    return (sp.kron(T, I) + sp.kron(I, T) + sp.kron(C, C) + 2.0 * sp.eye(side * side)).tocsr()


@pytest.mark.parametrize("mix", ["3_5_7_diagonals", "nine_diagonals"])
def test_batched_diagonal_storage_and_csr_fallback(mix):
    """Systems on at most 8 diagonals each are multiplied from batched
    diagonal storage (1D: 3, 2D: 5 and 3D: 7 diagonals in one batch); a
    system on 9 diagonals sends the batch to the CSR product. Every pair
    against FP64 Jacobi-PCG."""
    from mpsparse_amd import batched
    from mpsparse_amd.problems import diffusion_2d, diffusion_3d

    if mix == "3_5_7_diagonals":
        blocks = [_spd_1d(900, 1), diffusion_2d(30, sigma=1.0, seed=2).to_scipy(),
                  diffusion_3d(9, sigma=1.0, seed=3).to_scipy(), _spd_1d(70, 4)]
        planes = 8
    else:
        blocks = [diffusion_2d(24, sigma=1.0, seed=5).to_scipy(), _spd_9pt(20, 6)]
        planes = 0
    A, S, off, torch = _block_diag(blocks)
    n = int(off[-1])
    B = batched.batch_tensor(A, torch.from_numpy(off))
    b = np.random.default_rng(7).standard_normal((n, 4)).astype(np.float32)
    x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    say("batched storage", mix)
    info = B.pcg(torch.from_numpy(b).cuda(), x)
    assert info["diagonal_planes"] == planes, info["diagonal_planes"]
    assert info["converged_pairs"] == info["total_pairs"] == 4 * len(blocks)
    _check_pairs(S.tocsr(), off, b, x.cpu().numpy(), info, 1e-6, 4)


def test_batched_stencil_batch_uses_five_planes():
    import torch

    from mpsparse_amd import batched

    A, S, off, sides = _build(4, 10, 30, seed=9)
    B = batched.batch_tensor(A, torch.from_numpy(off))
    n = int(off[-1])
    b = np.random.default_rng(8).standard_normal((n, 4)).astype(np.float32)
    x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    say("batched five planes", sides.tolist())
    info = B.pcg(torch.from_numpy(b).cuda(), x)
    assert info["diagonal_planes"] == 5
    assert info["symmetric_storage"] == 1  # harmonic-mean faces: A(i,j) == A(j,i) bit for bit
    _check_pairs(S, off, b, x.cpu().numpy(), info, 1e-6, 4)


def test_batched_nonsymmetric_values_keep_five_planes():
    """The same kind of batch with one coupling of the last system changed
    (one side of it only): the product reads all five planes, and every pair
    still meets the FP64 check (Jacobi-PCG on the slightly nonsymmetric
    system converges here; its iteration bounds are not checked)."""
    import torch

    from gpu_util import csr_from_arrays
    from mpsparse_amd import batched

    A0, S0, off, sides = _build(3, 10, 30, seed=12)
    S = S0.tocsr().astype(np.float32)
    n = int(off[-1])
    row = n - 2
    e = S.indptr[row]  # the first (lower) coupling of that row
    S.data[e] *= 1.0 + 2.0 ** -10
    A = csr_from_arrays(S.indptr.astype(np.uint32), S.indices.astype(np.uint32), S.data, n, n)
    B = batched.batch_tensor(A, torch.from_numpy(off))
    b = np.random.default_rng(13).standard_normal((n, 4)).astype(np.float32)
    x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    say("batched nonsymmetric", sides.tolist())
    info = B.pcg(torch.from_numpy(b).cuda(), x)
    assert info["diagonal_planes"] == 5 and info["symmetric_storage"] == 0
    _check_pairs(S.astype(np.float64), off, b, x.cpu().numpy(), info, 1e-6, 4, its_check=False)
