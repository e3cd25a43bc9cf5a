"""Packed-key COO -> CSR oracle (test infrastructure only).

The reference converts ``u64`` keys ``(row<<32)|col`` with a stable 4-bit LSD
radix sort (iterative/cgd/coo_csr.metal:17-193, iterative/BiCGSTAB/spmv.cpp:62-228):
per-pass stable scatter with groups visited in index order, so the whole sort is
a *stable* sort of the keys -- duplicate keys keep their input order and are not
summed. The result is then compressed to ``row_ptr``/``col_ind`` and the values
are permuted alongside (coo_csr.metal:144-193).

Three independent formulations live here:

* :func:`coo_to_csr` -- numpy ``argsort(kind="stable")`` + ``searchsorted``
  (:func:`coo_to_csr_bucketed` is the same in parallel threads, for large inputs);
* :func:`coo_to_csr_lsd` -- the step-by-step C restatement of the reference's
  kernels (``lsd_coo_csr.c``, compiled by ``oracle/Makefile``);
* :func:`coo_to_csr_bruteforce` -- Python ``sorted`` over ``(key, index)``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class KeyRangeError(ValueError):
    """A key's row >= num_rows or col >= num_cols."""


def _check_range(keys: np.ndarray, num_rows: int, num_cols: int) -> None:
    if keys.size == 0:
        return
    rows = keys >> np.uint64(32)
    cols = keys & np.uint64(0xFFFFFFFF)
    if int(rows.max()) >= num_rows or int(cols.max()) >= num_cols:
        raise KeyRangeError("packed key outside [0,num_rows) x [0,num_cols)")


def coo_to_csr(keys, vals, num_rows: int, num_cols: int):
    """Stable-sort COO->CSR. Returns (row_ptr u32[n+1], col_ind u32[nnz], vals[nnz]).

    Semantics of iterative/BiCGSTAB/spmv.cpp:62-228 (sort) and
    coo_csr.metal:144-193 (compress + row fix): ``row_ptr[r]`` = index of the
    first sorted key with row >= r, ``row_ptr[n] = nnz``.
    """
    keys = np.ascontiguousarray(keys).view(np.uint64)
    vals = np.ascontiguousarray(vals)
    _check_range(keys, num_rows, num_cols)
    perm = np.argsort(keys, kind="stable")
    sk = keys[perm]
    rows = (sk >> np.uint64(32)).astype(np.int64)
    row_ptr = np.searchsorted(rows, np.arange(num_rows + 1, dtype=np.int64), side="left")
    col_ind = (sk & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return row_ptr.astype(np.uint32), col_ind, vals[perm]


def coo_to_csr_bucketed(keys, vals, num_rows: int, num_cols: int, workers: int = 16, buckets: int = 128):
    """:func:`coo_to_csr` for large inputs, in parallel threads; same result.

    The keys are split into `buckets` ranges of their row (the high half of
    the key, so every key of bucket b precedes every key of bucket b+1); each
    bucket keeps its keys in input order (a boolean mask preserves order) and
    is sorted with numpy's stable argsort in a thread pool (numpy releases the
    interpreter lock while sorting). Concatenating the buckets is therefore the
    stable sort of all keys. Pinned against :func:`coo_to_csr` in the tests."""
    from concurrent.futures import ThreadPoolExecutor

    keys = np.ascontiguousarray(keys).view(np.uint64)
    vals = np.ascontiguousarray(vals)
    _check_range(keys, num_rows, num_cols)
    nnz = keys.size
    if nnz == 0 or num_rows <= 1:
        return coo_to_csr(keys, vals, num_rows, num_cols)
    shift = max(0, int(num_rows - 1).bit_length() - max(1, int(buckets - 1).bit_length()))
    bid = ((keys >> np.uint64(32)) >> np.uint64(shift)).astype(np.uint8)  # buckets <= 256
    nb = int(bid.max()) + 1
    counts = np.bincount(bid, minlength=nb)
    order = np.argsort(bid, kind="stable")  # by bucket, input order inside (radix sort for 8-bit keys)
    starts = np.concatenate(([0], np.cumsum(counts)))
    perm = np.empty(nnz, dtype=np.int64)

    def one(b):
        idx = order[starts[b]:starts[b + 1]]
        perm[starts[b]:starts[b + 1]] = idx[np.argsort(keys[idx], kind="stable")]

    with ThreadPoolExecutor(max(1, workers)) as ex:
        list(ex.map(one, range(nb)))
    sk = keys[perm]
    rows = (sk >> np.uint64(32)).astype(np.int64)
    row_ptr = np.searchsorted(rows, np.arange(num_rows + 1, dtype=np.int64), side="left")
    col_ind = (sk & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return row_ptr.astype(np.uint32), col_ind, vals[perm]


def coo_to_csr_bruteforce(keys, vals, num_rows: int, num_cols: int):
    """Pure-Python restatement for tiny inputs: sort (key, position) pairs."""
    keys = [int(k) for k in np.asarray(keys).view(np.uint64)]
    vals = list(np.asarray(vals))
    for k in keys:
        if (k >> 32) >= num_rows or (k & 0xFFFFFFFF) >= num_cols:
            raise KeyRangeError("packed key outside the matrix")
    order = sorted(range(len(keys)), key=lambda i: (keys[i], i))
    row_ptr = [0] * (num_rows + 1)
    for i in order:
        row_ptr[(keys[i] >> 32) + 1] += 1
    for r in range(num_rows):
        row_ptr[r + 1] += row_ptr[r]
    col_ind = [keys[i] & 0xFFFFFFFF for i in order]
    out_vals = [vals[i] for i in order]
    return (np.array(row_ptr, dtype=np.uint32), np.array(col_ind, dtype=np.uint32),
            np.array(out_vals, dtype=np.asarray(vals).dtype if len(vals) else np.float32))


def build_lib(force: bool = False) -> str:
    """Compile ``lsd_coo_csr.c`` with the committed Makefile; returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        build_lib()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_lsd_coo_to_csr.restype = ctypes.c_int
        lib.oracle_lsd_coo_to_csr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_spmv_f32_seq.restype = None
        lib.oracle_spmv_f32_seq.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint32,
                                                                    ctypes.c_void_p, ctypes.c_void_p]
        _lib = lib
    return _lib


def coo_to_csr_lsd(keys, vals, num_rows: int, num_cols: int):
    """The reference's 4-bit LSD radix sort + compress + row fixer, restated in C."""
    lib = _load()
    keys = np.ascontiguousarray(keys).view(np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    nnz = keys.shape[0]
    row_ptr = np.empty(num_rows + 1, dtype=np.uint32)
    col_ind = np.empty(nnz, dtype=np.uint32)
    out_vals = np.empty(nnz, dtype=np.float32)
    rc = lib.oracle_lsd_coo_to_csr(keys.ctypes.data, vals.ctypes.data, nnz, num_rows, num_cols,
                                   row_ptr.ctypes.data, col_ind.ctypes.data, out_vals.ctypes.data)
    if rc == -2:
        raise KeyRangeError("packed key outside [0,num_rows) x [0,num_cols)")
    if rc != 0:
        raise MemoryError("oracle LSD sort: allocation failed")
    return row_ptr, col_ind, out_vals


def spmv_f32_seq(row_ptr, col_ind, vals, x):
    """Left-to-right FP32 row sums (C), see ``oracle_spmv_f32_seq``."""
    lib = _load()
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col_ind = np.ascontiguousarray(col_ind, dtype=np.uint32)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = row_ptr.shape[0] - 1
    y = np.empty(n, dtype=np.float32)
    lib.oracle_spmv_f32_seq(row_ptr.ctypes.data, col_ind.ctypes.data, vals.ctypes.data, n,
                            x.ctypes.data, y.ctypes.data)
    return y
