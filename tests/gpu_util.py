"""Helpers shared by the GPU tests (not collected by pytest)."""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    """Name the case before a GPU step (flushed, so it survives a crash)."""
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def t_keys(keys, device="cuda"):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(keys).view(np.int64))
    return t.to(device) if device != "cpu" else t


def t_f32(a, device="cuda"):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    return t.to(device) if device != "cpu" else t


def build_csr(keys, vals, n, m, device="cuda", export=True):
    """spmv.csr_tensor from numpy COO; returns (A, row_ptr, col_ind, vals) as numpy."""
    import torch

    from mpsparse_amd import spmv

    nnz = keys.shape[0]
    rp = torch.zeros(n + 1, dtype=torch.int32, device=device)
    ci = torch.zeros(nnz, dtype=torch.int32, device=device)
    ov = torch.zeros(nnz, dtype=torch.float32, device=device)
    A = spmv.csr_tensor(t_keys(keys, device), t_f32(vals, device), rp if export else torch.empty(0),
                        ci if export else torch.empty(0), ov if export else torch.empty(0), n, m)
    return A, rp.cpu().numpy().view(np.uint32), ci.cpu().numpy().view(np.uint32), ov.cpu().numpy()


def csr_from_arrays(row_ptr, col_ind, vals, n, m):
    """A handle built from sorted CSR arrays (via packed keys, order preserved)."""
    from mpsparse_amd.problems import pack_keys

    rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(np.asarray(row_ptr, dtype=np.int64)))
    keys = pack_keys(rows, np.asarray(col_ind, dtype=np.uint64))
    return build_csr(keys, vals, n, m, export=False)[0]


class Lib:
    """Direct ctypes view of the C ABI (same library the extension links)."""

    def __init__(self):
        self.lib = ctypes.CDLL(os.path.join(ROOT, "mpsparse_amd", "libmpsparse_amd.so"))
        self.lib.msp_last_error.restype = ctypes.c_char_p

    def stream(self):
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def fp32_floor(A, b):
    """Relative residual of the exact solution rounded to FP32: no FP32-stored x can do better."""
    import scipy.sparse.linalg as spla

    from oracle import krylov as K

    xs = spla.spsolve(A.tocsc(), np.asarray(b, dtype=np.float64))
    return K.true_relres(A, xs.astype(np.float32), b)


def check_fp32_solution(info, A, x, b, tol):
    """A converged FP32 solve meets tol in the FP64 true residual; a stagnated one
    is accepted only when the FP32 rounding floor is within 3x of tol, and then
    must lie within 4x of that floor."""
    from oracle import krylov as K

    rel = K.true_relres(A, x, b)
    if info["status"] == "converged":
        assert rel <= 1.05 * tol, (rel, info)
        return rel
    floor = fp32_floor(A, b)
    assert info["status"] == "stagnated" and floor > tol / 3, (info, floor, rel)
    assert rel <= max(1.05 * tol, 4 * floor), (info, floor, rel)
    return rel
