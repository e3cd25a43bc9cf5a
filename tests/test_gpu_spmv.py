"""GPU CSR SpMV against the FP64 oracle, within the FP32 forward error bound.

Tolerance: |y - y64| <= 2 * gamma_k (|A||x|)_i + 1e-30 per row, gamma_k with
k = row length + 1 and u = 2^-24 (oracle.spmv.spmv_error_bound); the factor 2
covers the different (tree) summation order. BF16/FP16 value storage is checked
against the FP64 product of the rounded values with the same bound.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import Lib, build_csr, csr_from_arrays, say, t_f32
from mpsparse_amd.problems import block_sparse, diffusion_3d, laplacian_2d, pack_keys
from oracle import coo_csr as oc
from oracle import spmv as S

pytestmark = pytest.mark.gpu


def _check(A, rp, ci, v, m, rng, values="f32", vals_used=None, dlo=None):
    import torch

    x = rng.standard_normal(m).astype(np.float32)
    y = torch.full((A.num_rows,), np.nan, dtype=torch.float32, device="cuda")
    A.mv(t_f32(x), y, values=values)
    y = y.cpu().numpy()
    vu = v if vals_used is None else vals_used
    y64 = S.csr_spmv(rp, ci, vu, x)
    bound = 2 * S.spmv_error_bound(rp, ci, vu, x) + 1e-30
    if dlo is not None:  # 16-bit values: the FP32 diagonal remainder is one more term
        y64 = y64 + dlo * x[:dlo.size]
        bound = bound + 4 * 2.0 ** -24 * np.abs(dlo * x[:dlo.size])
    err = np.abs(y.astype(np.float64) - y64)
    assert np.all(np.isfinite(y))
    assert np.all(err <= bound), (err.max(), np.argmax(err - bound))


def test_laplacian():
    p = laplacian_2d(64)
    say("spmv laplacian")
    A, rp, ci, v = build_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    _check(A, rp, ci, v, p.num_cols, np.random.default_rng(0))


def test_block_sparse_144_per_row():
    q = block_sparse(1024, 16, 8, seed=1)
    say("spmv block sparse 1024")
    A, rp, ci, v = build_csr(q.keys, q.vals, q.num_rows, q.num_cols)
    _check(A, rp, ci, v, q.num_cols, np.random.default_rng(1))


def test_diffusion_3d():
    M = diffusion_3d(40, sigma=1.0, seed=1)
    say("spmv diffusion 40^3")
    A = csr_from_arrays(M.row_ptr, M.col_ind, M.vals, M.num_rows, M.num_cols)
    _check(A, M.row_ptr, M.col_ind, M.vals, M.num_cols, np.random.default_rng(2))


@pytest.mark.parametrize("shape", ["ragged", "long_rows", "nonsquare", "empty_tail", "single_row"])
def test_irregular(shape):
    rng = np.random.default_rng(hash(shape) % 1000)
    if shape == "ragged":
        n, m = 20000, 20000
        lens = rng.integers(0, 300, n) * (rng.random(n) < 0.3)
    elif shape == "long_rows":  # rows longer than one LDS window (4096)
        n, m = 3000, 200000
        lens = rng.integers(0, 20, n)
        lens[[0, 17, 1500, 2999]] = [9000, 20000, 4097, 150000]
    elif shape == "nonsquare":
        n, m = 1000, 300000
        lens = rng.integers(1, 60, n)
    elif shape == "empty_tail":
        n, m = 50000, 50000
        lens = np.zeros(n, dtype=np.int64)
        lens[:100] = rng.integers(1, 30, 100)
    else:
        n, m = 1, 100000
        lens = np.array([70000])
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, m, rows.size)
    keys = pack_keys(rows, cols)
    vals = rng.standard_normal(rows.size).astype(np.float32)
    say("spmv irregular", shape, n, m, rows.size)
    A, rp, ci, v = build_csr(keys, vals, n, m)
    _check(A, rp, ci, v, m, rng)


@pytest.mark.parametrize("values", ["bf16", "f16"])
def test_reduced_precision_values(values):
    M = diffusion_3d(32, sigma=1.0, seed=3)
    say("spmv", values)
    A = csr_from_arrays(M.row_ptr, M.col_ind, M.vals, M.num_rows, M.num_cols)
    used, dlo = S.round16_rowsum(M.row_ptr, M.col_ind, M.vals, values)
    _check(A, M.row_ptr, M.col_ind, M.vals, M.num_cols, np.random.default_rng(4), values=values,
           vals_used=used, dlo=dlo)


def test_direct_c_abi_call():
    """Same SpMV through ctypes on the C ABI with device pointers."""
    import torch

    p = laplacian_2d(64)
    A, rp, ci, v = build_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    L = Lib()
    x = torch.randn(p.num_cols, device="cuda")
    y = torch.zeros(p.num_rows, device="cuda")
    say("c-abi spmv")
    st = L.lib.msp_csr_spmv(ctypes.c_void_p(A.handle), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                            0, L.stream())
    assert st == 0, L.lib.msp_last_error()
    torch.cuda.synchronize()
    y64 = S.csr_spmv(rp, ci, v, x.cpu().numpy())
    assert np.allclose(y.cpu().numpy(), y64, atol=1e-5)


def test_cpu_tensors_round_trip():
    import torch

    p = laplacian_2d(32)
    A, rp, ci, v = build_csr(p.keys, p.vals, p.num_rows, p.num_cols, device="cpu")
    x = torch.randn(p.num_cols)
    y = torch.zeros(p.num_rows)
    say("cpu tensors mv")
    A.mv(x, y)
    assert np.allclose(y.numpy(), S.csr_spmv(rp, ci, v, x.numpy()), atol=1e-5)
    r2, c2, v2 = oc.coo_to_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    assert np.array_equal(rp, r2) and np.array_equal(ci, c2) and np.array_equal(v, v2)
