"""msp_coo_to_csr_noalloc (temporaries in a fixed per-device scratch array, no
allocation) against msp_coo_to_csr, through the C ABI with device pointers.

Bar: bit-exact row_ptr / col_ind / values (integer and permutation work), on
consecutive calls whose inputs shrink (the scratch keeps the previous call's
contents), on the single-level fallback (a row too long for one segment) at
the size limit, and refusals of over-limit sizes and host pointers.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import Lib, say
from mpsparse_amd.problems import pack_keys
from oracle import coo_csr as oc

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20


def _keys(n, m, nnz, seed, heavy=0):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, nnz)
    if heavy:
        rows[:heavy] = n // 2  # one row longer than a segment holds
    cols = rng.integers(0, m, nnz)
    cols[: nnz // 10] = cols[nnz // 10: 2 * (nnz // 10)]  # duplicate keys, kept in input order
    rows[: nnz // 10] = rows[nnz // 10: 2 * (nnz // 10)]
    perm = rng.permutation(nnz)
    return pack_keys(rows[perm], cols[perm]), rng.standard_normal(nnz).astype(np.float32)


def _run(L, fn, keys, vals, n, m):
    import torch

    nnz = keys.size
    k = torch.from_numpy(keys.view(np.int64)).cuda()
    v = torch.from_numpy(vals).cuda()
    rp = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
    ci = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda")
    ov = torch.full((max(nnz, 1),), np.nan, dtype=torch.float32, device="cuda")
    st = getattr(L.lib, fn)(ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(v.data_ptr()), ctypes.c_uint64(nnz),
                            ctypes.c_uint32(n), ctypes.c_uint32(m), ctypes.c_void_p(rp.data_ptr()),
                            ctypes.c_void_p(ci.data_ptr()), ctypes.c_void_p(ov.data_ptr()), L.stream())
    assert st == 0, (fn, L.lib.msp_last_error())
    torch.cuda.synchronize()
    return rp.cpu().numpy().view(np.uint32), ci.cpu().numpy()[:nnz].view(np.uint32), ov.cpu().numpy()[:nnz]


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


@pytest.mark.parametrize("order", ["large_then_small", "limit_then_fallback"])
def test_noalloc_matches_allocating_entry_point(order):
    L = Lib()
    if order == "large_then_small":
        cases = [(50_000, 60_000, 300_000, 1, 0), (3_000, 70, 20_000, 2, 0)]
    else:  # 2^20 keys of 40-bit compacted keys, then the fallback path at the limit
        cases = [(1 << 20, 1 << 20, LIMIT, 3, 0), (200_000, 1 << 20, LIMIT - 7, 4, 10_000)]
    for n, m, nnz, seed, heavy in cases:
        keys, vals = _keys(n, m, nnz, seed, heavy)
        say("noalloc", order, n, m, nnz, heavy)
        got = _run(L, "msp_coo_to_csr_noalloc", keys, vals, n, m)
        ref = _run(L, "msp_coo_to_csr", keys, vals, n, m)
        _same(got, ref)
        _same(got, oc.coo_to_csr(keys, vals, n, m))


def test_noalloc_refuses_what_it_cannot_do():
    import torch

    L = Lib()
    say("noalloc refusals")
    big = torch.zeros(LIMIT + 1, dtype=torch.int64, device="cuda")
    bv = torch.zeros(LIMIT + 1, dtype=torch.float32, device="cuda")
    rp = torch.zeros(2, dtype=torch.int32, device="cuda")
    ci = torch.zeros(LIMIT + 1, dtype=torch.int32, device="cuda")
    st = L.lib.msp_coo_to_csr_noalloc(ctypes.c_void_p(big.data_ptr()), ctypes.c_void_p(bv.data_ptr()),
                                      ctypes.c_uint64(LIMIT + 1), ctypes.c_uint32(1), ctypes.c_uint32(1),
                                      ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(ci.data_ptr()),
                                      ctypes.c_void_p(bv.data_ptr()), L.stream())
    assert st == 1 and b"limit" in L.lib.msp_last_error()
    hk = np.zeros(4, dtype=np.uint64)
    hv = np.zeros(4, dtype=np.float32)
    st = L.lib.msp_coo_to_csr_noalloc(hk.ctypes.data_as(ctypes.c_void_p), hv.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_uint64(4), ctypes.c_uint32(1), ctypes.c_uint32(1),
                                      ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(ci.data_ptr()),
                                      ctypes.c_void_p(bv.data_ptr()), L.stream())
    assert st == 1 and b"device memory" in L.lib.msp_last_error()
