"""On-disk sparse matrices for real-matrix validation: Matrix Market coordinate
files and scipy-compatible CSR ``.npz`` archives, read into the packed-key COO
form every constructor of this package takes.

The reference has no file format (its inputs are in-memory packed keys,
README.md:34-42); these readers produce exactly those keys, so
``spmv.csr_tensor(*to_csr_tensor_args(read_matrix_market(path)))`` builds a
matrix from a file.
"""
from __future__ import annotations

import gzip
import io as _io

import numpy as np

from .problems import pack_keys

__all__ = ["read_matrix_market", "write_matrix_market", "load_npz", "save_npz", "to_csr_tensor_args"]


def _open(path, mode):
    if str(path).endswith(".gz"):
        return gzip.open(path, mode + "t")
    return open(path, mode)


def read_matrix_market(path):
    """Read a Matrix Market ``coordinate`` file (real, integer or pattern;
    general, symmetric or skew-symmetric). Returns (keys u64[nnz], vals f32[nnz],
    num_rows, num_cols) with 0-based packed keys ``(row << 32) | col`` in file
    order; symmetric files are expanded to both triangles (mirror entries follow
    the stored ones), pattern files get value 1."""
    with _open(path, "r") as f:
        header = f.readline().strip().split()
        if len(header) < 5 or header[0].lower() != "%%matrixmarket" or header[1].lower() != "matrix":
            raise ValueError(f"{path}: not a Matrix Market matrix file")
        fmt, field, sym = header[2].lower(), header[3].lower(), header[4].lower()
        if fmt != "coordinate":
            raise ValueError(f"{path}: only the coordinate format is supported (got {fmt})")
        if field not in ("real", "integer", "pattern", "double"):
            raise ValueError(f"{path}: field {field} is not supported (real, integer or pattern)")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise ValueError(f"{path}: symmetry {sym} is not supported")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        n, m, nnz = (int(t) for t in line.split()[:3])
        cols_per = 2 if field == "pattern" else 3
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nnz) if nnz else np.zeros((0, cols_per))
    if data.shape[0] != nnz or data.shape[1] < cols_per:
        raise ValueError(f"{path}: expected {nnz} entries of {cols_per} fields")
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = np.ones(nnz) if field == "pattern" else data[:, 2]
    if nnz and (r.min() < 0 or c.min() < 0 or r.max() >= n or c.max() >= m):
        raise ValueError(f"{path}: an index lies outside the {n} x {m} matrix")
    if sym != "general":
        off = r != c
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, sign * v[off]])
    keys = pack_keys(r.astype(np.uint64), c.astype(np.uint64))
    return keys, v.astype(np.float32), n, m


def write_matrix_market(path, row_ptr, col_ind, vals, num_rows, num_cols, comment=None):
    """Write a CSR matrix as a ``coordinate real general`` Matrix Market file
    (values with 9 significant digits: FP32 round-trips exactly)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    rows = np.repeat(np.arange(num_rows, dtype=np.int64), np.diff(row_ptr))
    cols = np.asarray(col_ind, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float32)
    buf = _io.StringIO()
    buf.write("%%MatrixMarket matrix coordinate real general\n")
    if comment:
        for ln in str(comment).splitlines():
            buf.write(f"% {ln}\n")
    buf.write(f"{num_rows} {num_cols} {rows.size}\n")
    np.savetxt(buf, np.column_stack([rows + 1, cols + 1, vals.astype(np.float64)]), fmt=["%d", "%d", "%.9g"])
    with _open(path, "w") as f:
        f.write(buf.getvalue())


def load_npz(path):
    """Read a scipy ``save_npz`` archive of a CSR/CSC/COO matrix (arrays only,
    no pickles). Returns (keys, vals, num_rows, num_cols) in row-major order."""
    with np.load(path, allow_pickle=False) as z:
        fmt = z["format"].item()
        fmt = fmt.decode() if isinstance(fmt, bytes) else str(fmt)
        n, m = (int(t) for t in z["shape"])
        if fmt == "csr":
            indptr, indices, data = z["indptr"], z["indices"], z["data"]
            rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr.astype(np.int64)))
            cols = indices.astype(np.int64)
        elif fmt == "csc":
            indptr, indices, data = z["indptr"], z["indices"], z["data"]
            cols = np.repeat(np.arange(m, dtype=np.int64), np.diff(indptr.astype(np.int64)))
            rows = indices.astype(np.int64)
        elif fmt == "coo":
            rows, cols, data = z["row"].astype(np.int64), z["col"].astype(np.int64), z["data"]
        else:
            raise ValueError(f"{path}: format {fmt} is not supported (csr, csc or coo)")
    if np.iscomplexobj(data):
        raise ValueError(f"{path}: complex values are not supported")
    keys = pack_keys(rows.astype(np.uint64), cols.astype(np.uint64))
    return keys, np.asarray(data, dtype=np.float32), n, m


def save_npz(path, row_ptr, col_ind, vals, num_rows, num_cols):
    """Write a CSR matrix in scipy's ``save_npz`` layout (uncompressed)."""
    np.savez(path, indptr=np.asarray(row_ptr, dtype=np.int32), indices=np.asarray(col_ind, dtype=np.int32),
             data=np.asarray(vals, dtype=np.float32), shape=np.array([num_rows, num_cols], dtype=np.int64),
             format=np.array("csr"))


def to_csr_tensor_args(matrix, device="cuda"):
    """(keys, values, row_ptr, col_ind, out_vals, n, m) for spmv.csr_tensor from
    a (keys, vals, n, m) tuple; the CSR buffers are allocated so that the
    constructor fills them."""
    import torch

    keys, vals, n, m = matrix
    k = torch.from_numpy(np.ascontiguousarray(keys).view(np.int64)).to(device)
    v = torch.from_numpy(np.ascontiguousarray(vals, dtype=np.float32)).to(device)
    rp = torch.empty(n + 1, dtype=torch.int32, device=device)
    ci = torch.empty(k.numel(), dtype=torch.int32, device=device)
    ov = torch.empty(k.numel(), dtype=torch.float32, device=device)
    return k, v, rp, ci, ov, n, m
