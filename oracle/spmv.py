"""SpMV oracle (test infrastructure only).

* :func:`csr_spmv` -- ``y = A x`` in FP64 (or any numpy dtype) by an explicit
  segmented sum, the arithmetic of the reference's ``spmv_op``
  (iterative/BiCGSTAB/spmv.metal:11-49) without its 32-lane tree order.
* :func:`bsr_spmv` -- block-row SpMV over row-major ``[nblk][b][b]`` blocks, the
  layout of the reference's BCSR kernels (iterative/BiCGSTAB/bcsr_spmv.metal:30-69)
  and of scipy's ``tobsr`` data it is fed from (iterative/BiCGSTAB/bcsr_benchmark.py:36-66).
* :func:`csr_to_bsr` -- CSR -> BCSR arrays (block pattern, dense zero-filled
  blocks, duplicates summed), an explicit restatement checked against scipy.
* :func:`spmv_error_bound` -- the standard forward bound |y^ - y| <= gamma_k |A||x|
  with k the row length, used to size FP32 tolerances.
* :func:`round16_rowsum` -- the library's 16-bit value copy (BF16 / FP16
  off-diagonals, diagonal absorbing their rounding defects as a 16-bit slot
  value plus an FP32 remainder). Pinned by row sums equal to A's, off-diagonals
  equal to plain rounding, and, for diffusion M-matrices, the energy bound
  |x^T (A16 - A) x| <= u x^T A x with u the format's unit roundoff.
"""
from __future__ import annotations

import numpy as np


def csr_spmv(row_ptr, col_ind, vals, x, dtype=np.float64):
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col_ind = np.asarray(col_ind, dtype=np.int64)
    prod = np.asarray(vals, dtype=dtype) * np.asarray(x, dtype=dtype)[col_ind]
    n = row_ptr.shape[0] - 1
    y = np.zeros(n, dtype=dtype)
    counts = np.diff(row_ptr)
    nz = counts > 0
    if prod.size:
        sums = np.add.reduceat(prod, row_ptr[:-1][nz])
        y[nz] = sums
    return y


def spmv_error_bound(row_ptr, col_ind, vals, x, unit_roundoff=2.0 ** -24):
    """Per-row bound gamma_k * (|A| |x|)_i, k = row length (+1 for the product)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    k = np.diff(row_ptr).astype(np.float64) + 1.0
    gamma = k * unit_roundoff / (1.0 - k * unit_roundoff)
    absy = csr_spmv(row_ptr, col_ind, np.abs(np.asarray(vals, dtype=np.float64)),
                    np.abs(np.asarray(x, dtype=np.float64)))
    return gamma * absy


def csr_to_bsr(row_ptr, col_ind, vals, num_rows, num_cols, block):
    """CSR -> BCSR: (brow_ptr u32[nbr+1], bcol u32[nblk], blocks f32[nblk,b,b], nbr, nbc).

    Rows/cols are padded up to a multiple of ``block``; blocks within a block-row
    are ordered by block column; entries falling in the same (row, col) are summed
    in storage order (as scipy ``tobsr`` does for duplicate COO entries).
    """
    B = int(block)
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col_ind = np.asarray(col_ind, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float32)
    nbr = (num_rows + B - 1) // B
    nbc = (num_cols + B - 1) // B
    rows = np.repeat(np.arange(num_rows, dtype=np.int64), np.diff(row_ptr))
    bkey = (rows // B) * nbc + col_ind // B
    uniq, inv = np.unique(bkey, return_inverse=True)
    nblk = uniq.shape[0]
    blocks = np.zeros((nblk, B, B), dtype=np.float64)
    np.add.at(blocks, (inv, rows % B, col_ind % B), vals.astype(np.float64))
    brow = uniq // nbc
    bcol = (uniq % nbc).astype(np.uint32)
    brow_ptr = np.searchsorted(brow, np.arange(nbr + 1), side="left").astype(np.uint32)
    return brow_ptr, bcol, blocks.astype(np.float32), nbr, nbc


def bsr_spmv(brow_ptr, bcol, blocks, x, dtype=np.float64):
    """y = A x for BCSR with row-major blocks; x is padded to nbc*b."""
    brow_ptr = np.asarray(brow_ptr, dtype=np.int64)
    bcol = np.asarray(bcol, dtype=np.int64)
    blocks = np.asarray(blocks, dtype=dtype)
    B = blocks.shape[1]
    nbr = brow_ptr.shape[0] - 1
    x = np.asarray(x, dtype=dtype)
    xb = x.reshape(-1, B)
    prod = np.einsum("kij,kj->ki", blocks, xb[bcol]) if bcol.size else np.zeros((0, B), dtype)
    y = np.zeros((nbr, B), dtype=dtype)
    brows = np.repeat(np.arange(nbr), np.diff(brow_ptr))
    np.add.at(y, brows, prod)
    return y.reshape(-1)


def round_bf16(a):
    """Round-to-nearest-even FP32 -> BF16, widened back to float32.

    Finite values and infinities: add half an ulp of the BF16 result minus one,
    plus the lowest kept bit (ties go to the even neighbour), then truncate;
    values past the largest BF16 overflow to infinity. NaNs stay NaN (quietened,
    sign and high payload bits kept); adding the bias to a NaN's bits could
    carry into the sign and give -0. Pinned against torch's bfloat16 cast in
    tests/test_oracle_bf16.py."""
    u32 = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = u32.astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    nan = ((u32 & 0x7F800000) == 0x7F800000) & ((u32 & 0x007FFFFF) != 0)
    r = np.where(nan, (u32 >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).view(np.float32)


def round16_rowsum(row_ptr, col_ind, vals, fmt="bf16", num_cols=None):
    """(vals16 float64[nnz], dlo float64[n]): off-diagonals rounded to `fmt`,
    d_i = a_ii + sum_{j != i}(a_ij - round(a_ij)) (FP64) stored as round(d_i) in
    the first diagonal slot (other diagonal slots 0) and dlo_i = d_i - round(d_i)
    rounded to FP32 (0 for rows i >= num_cols, which have no diagonal column).
    y = vals16-SpMV + dlo * x is the library's 16-bit product."""
    rnd = round_bf16 if fmt == "bf16" else (lambda a: np.asarray(a, np.float32).astype(np.float16).astype(np.float32))
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col = np.asarray(col_ind, dtype=np.int64)
    v = np.asarray(vals, dtype=np.float32)
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    diag = col == rows
    out = rnd(v).astype(np.float64)
    defect = np.where(diag, v.astype(np.float64), v.astype(np.float64) - out)
    d = np.zeros(n)
    np.add.at(d, rows, defect)
    out[diag] = 0.0
    first = np.full(n, -1, dtype=np.int64)
    idx = np.nonzero(diag)[0]
    # first diagonal slot per row (indices ascend, so the minimum per row)
    first_rows = rows[idx]
    order = np.lexsort((idx, first_rows))
    idx, first_rows = idx[order], first_rows[order]
    keep = np.r_[True, first_rows[1:] != first_rows[:-1]] if idx.size else np.zeros(0, bool)
    first[first_rows[keep]] = idx[keep]
    hi = np.zeros(n)
    has = first >= 0
    hi[has] = rnd(d[has].astype(np.float32)).astype(np.float64)
    out[first[has]] = hi[has]
    dlo = (d - hi).astype(np.float32).astype(np.float64)
    if num_cols is not None:
        dlo[np.arange(n) >= num_cols] = 0.0
    return out, dlo
