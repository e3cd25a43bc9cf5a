"""Seeded synthetic sparse systems: the workloads the solver library is measured on.

Every generator is deterministic for a given seed and returns plain numpy arrays,
so the same inputs can be handed to the HIP path and to the CPU oracle.

COO inputs use the packed-key layout of the reference's harness
(``pack_keys``: ``(int64(row) << 32) | int64(col)``, README.md:34-42,
iterative/cgd/benchmark.py:112-115): one ``u64`` key per nonzero plus one
``float32`` value, in arbitrary (here: shuffled) order.

Workloads (sizes and distributions are fixed here):

* ``laplacian_2d(64)``      -- 2D 5-point Dirichlet Laplacian, 64x64 grid
  (n=4096, nnz=20224), keys shuffled with seed 0, b ~ N(0,1).
* ``block_sparse(65536, 16, 8)`` -- 65,536 block-rows of 16x16 blocks: the
  diagonal block plus 8 distinct uniform random off-diagonal block-columns,
  entries N(0,1)/sqrt(144), a_ii = sum_{j!=i}|a_ij| + 1 (nonsymmetric, strictly
  diagonally dominant). n = 1,048,576, nnz = 150,994,944. Seed 0.
* ``diffusion_3d(256, 256, 256, sigma=1, seed=1)`` -- 7-point finite-volume
  diffusion with log-normal cell conductivities, face coefficients the harmonic
  mean of the two cells, Dirichlet faces weighted 2*k (half-cell distance).
  SPD. n = 16,777,216, nnz = 117,047,296.
* ``diffusion_2d`` -- the 5-point 2D analogue (batched and training sets).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "pack_keys",
    "unpack_keys",
    "CooProblem",
    "CsrMatrix",
    "laplacian_2d",
    "block_sparse",
    "diffusion_2d",
    "diffusion_3d",
    "csr_to_coo_keys",
    "batched_diffusion_sides",
]


def pack_keys(rows, cols):
    """``(int64(row) << 32) | int64(col)`` as uint64 (reference README.md:34-42).

    Accepts numpy arrays or torch tensors; returns the same kind (torch keeps
    int64, the dtype the reference's ``csr_tensor`` takes, spmv.cpp:243).
    """
    try:
        import torch

        if isinstance(rows, torch.Tensor):
            return (rows.to(torch.int64) << 32) | cols.to(torch.int64)
    except ImportError:  # pragma: no cover - torch is always present here
        pass
    rows = np.asarray(rows).astype(np.uint64)
    cols = np.asarray(cols).astype(np.uint64)
    return (rows << np.uint64(32)) | cols


def unpack_keys(keys):
    """Inverse of :func:`pack_keys` -> (rows, cols) as uint32 (or int32 tensors)."""
    try:
        import torch

        if isinstance(keys, torch.Tensor):
            return (keys >> 32).to(torch.int32), (keys & 0xFFFFFFFF).to(torch.int32)
    except ImportError:  # pragma: no cover
        pass
    keys = np.asarray(keys).view(np.uint64)
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)


@dataclass
class CsrMatrix:
    """Sorted CSR with uint32 indices and float32 values (host arrays)."""

    row_ptr: np.ndarray  # uint32[n+1] (int64 when nnz >= 2**32 would be needed)
    col_ind: np.ndarray  # uint32[nnz]
    vals: np.ndarray  # float32[nnz]
    num_rows: int
    num_cols: int
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.col_ind.shape[0])

    def to_scipy(self, dtype=np.float64):
        import scipy.sparse as sp

        return sp.csr_matrix(
            (self.vals.astype(dtype), self.col_ind.astype(np.int64), self.row_ptr.astype(np.int64)),
            shape=(self.num_rows, self.num_cols),
        )

    def rows(self) -> np.ndarray:
        """Row index of every nonzero (uint32[nnz])."""
        counts = np.diff(self.row_ptr.astype(np.int64))
        return np.repeat(np.arange(self.num_rows, dtype=np.uint32), counts)


@dataclass
class CooProblem:
    """A linear system given as shuffled packed COO keys plus a right-hand side."""

    keys: np.ndarray  # uint64[nnz]
    vals: np.ndarray  # float32[nnz]
    num_rows: int
    num_cols: int
    b: np.ndarray  # float32[num_rows]
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.keys.shape[0])


def csr_to_coo_keys(A: CsrMatrix, rng: np.random.Generator | None = None):
    """Packed keys + values of a CSR matrix, shuffled by ``rng`` when given."""
    keys = pack_keys(A.rows(), A.col_ind)
    vals = A.vals
    if rng is not None:
        perm = rng.permutation(keys.shape[0])
        keys = keys[perm]
        vals = vals[perm]
    return keys, vals


def _stencil_to_csr(cols: np.ndarray, vals: np.ndarray, n: int, meta: dict) -> CsrMatrix:
    """(n, k) column/value tables with -1 marking absent entries -> sorted CSR."""
    mask = cols >= 0
    counts = mask.sum(axis=1)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col_ind = cols[mask].astype(np.uint32)
    v = vals[mask].astype(np.float32)
    return CsrMatrix(row_ptr.astype(np.uint32), col_ind, v, n, n, meta)


def laplacian_2d(grid: int, seed: int = 0) -> CooProblem:
    """2D 5-point Dirichlet Laplacian (diag 4, neighbours -1), shuffled packed keys.

    Row/column index of grid point (i, j) is ``i*grid + j``. The nonzeros are
    shuffled with ``default_rng(seed).permutation`` and then
    ``b = default_rng(seed)`` (same generator, next draws) ``.standard_normal(n)``.
    """
    n = grid * grid
    idx = np.arange(n, dtype=np.int64).reshape(grid, grid)
    cols = np.full((n, 5), -1, dtype=np.int64)
    vals = np.zeros((n, 5), dtype=np.float64)
    # column order within a row: -i, -j, diag, +j, +i
    cols[grid:, 0] = (idx[1:, :] - grid).ravel()
    vals[grid:, 0] = -1.0
    c = np.full((grid, grid), -1, dtype=np.int64)
    c[:, 1:] = idx[:, 1:] - 1
    cols[:, 1] = c.ravel()
    vals[:, 1] = np.where(cols[:, 1] >= 0, -1.0, 0.0)
    cols[:, 2] = idx.ravel()
    vals[:, 2] = 4.0
    c = np.full((grid, grid), -1, dtype=np.int64)
    c[:, :-1] = idx[:, :-1] + 1
    cols[:, 3] = c.ravel()
    vals[:, 3] = np.where(cols[:, 3] >= 0, -1.0, 0.0)
    cols[: n - grid, 4] = (idx[:-1, :] + grid).ravel()
    vals[: n - grid, 4] = -1.0
    A = _stencil_to_csr(cols, vals, n, {"kind": "laplacian_2d", "grid": grid})
    rng = np.random.default_rng(seed)
    keys, v = csr_to_coo_keys(A, rng)
    b = rng.standard_normal(n).astype(np.float32)
    return CooProblem(keys, v, n, n, b, {"kind": "laplacian_2d", "grid": grid, "seed": seed})


def laplacian_2d_csr(grid: int) -> CsrMatrix:
    """The sorted CSR form of :func:`laplacian_2d` (no shuffle)."""
    p = laplacian_2d(grid)
    order = np.argsort(p.keys, kind="stable")
    k = p.keys[order]
    rows = (k >> np.uint64(32)).astype(np.int64)
    row_ptr = np.searchsorted(rows, np.arange(p.num_rows + 1), side="left")
    return CsrMatrix(row_ptr.astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32),
                     p.vals[order], p.num_rows, p.num_cols, dict(p.meta))


def block_sparse(num_block_rows: int = 65536, block: int = 16, off_blocks: int = 8,
                 seed: int = 0, shuffle: bool = True) -> CooProblem:
    """Nonsymmetric, strictly diagonally dominant block-sparse system.

    Each block-row holds its diagonal block plus ``off_blocks`` distinct
    off-diagonal block-columns drawn uniformly. Entries are N(0,1)/sqrt(144)
    (float32), then each diagonal entry becomes ``sum_{j!=i}|a_ij| + 1``
    (summed in float64, rounded once to float32). Keys are shuffled with a
    uniform random permutation; then b ~ N(0,1).
    """
    nbr = num_block_rows
    B = block
    nb = off_blocks + 1
    rng = np.random.default_rng(seed)
    br = np.arange(nbr, dtype=np.int64)
    if off_blocks > nbr - 1:
        raise ValueError("more off-diagonal blocks than block columns")
    bc = rng.integers(0, nbr - 1, size=(nbr, off_blocks), dtype=np.int64)
    bc += bc >= br[:, None]
    while True:
        s = np.sort(bc, axis=1)
        dup = (s[:, 1:] == s[:, :-1]).any(axis=1)
        if not dup.any():
            break
        rows_bad = np.nonzero(dup)[0]
        redraw = rng.integers(0, nbr - 1, size=(rows_bad.size, off_blocks), dtype=np.int64)
        redraw += redraw >= br[rows_bad, None]
        bc[rows_bad] = redraw
    bcols = np.concatenate([br[:, None], bc], axis=1)  # (nbr, nb); diagonal block first
    vals = rng.standard_normal((nbr, nb, B, B), dtype=np.float32)
    vals /= np.float32(np.sqrt(144.0))
    ii = np.arange(B)
    absrow = np.abs(vals).astype(np.float64).sum(axis=(1, 3))  # (nbr, B)
    absrow -= np.abs(vals[:, 0, ii, ii]).astype(np.float64)
    vals[:, 0, ii, ii] = (absrow + 1.0).astype(np.float32)
    rows = (br[:, None] * B + ii[None, :]).astype(np.uint64)  # (nbr, B)
    cols = (bcols[:, :, None] * B + ii[None, None, :]).astype(np.uint64)  # (nbr, nb, B)
    keys = (rows[:, None, :, None] << np.uint64(32)) | cols[:, :, None, :]
    keys = keys.reshape(-1)
    vals = vals.reshape(-1)
    if shuffle:
        perm = rng.permutation(keys.shape[0])
        keys = keys[perm]
        vals = vals[perm]
    n = nbr * B
    b = rng.standard_normal(n).astype(np.float32)
    meta = {"kind": "block_sparse", "num_block_rows": nbr, "block": B,
            "off_blocks": off_blocks, "seed": seed}
    return CooProblem(keys, vals, n, n, b, meta)


def _lognormal_k(shape, sigma: float, rng: np.random.Generator) -> np.ndarray:
    return np.exp(sigma * rng.standard_normal(shape))


def _harmonic(a, b):
    return 2.0 * a * b / (a + b)


def diffusion_stencil(k, new_int, new_float, arange):
    """(cols[n, 7], vals[n, 7]) of the finite-volume diffusion operator for a
    (nz, ny, nx) conductivity array ``k`` (FP64). Interior face coefficient:
    harmonic mean of the two cells; a Dirichlet boundary face adds ``2*k`` to
    the diagonal; a single plane (nz = 1) gives the 2D 5-point operator. Slots
    are (-z, -y, -x, diag, +x, +y, +z); absent neighbours have column -1.
    Works on numpy arrays or torch tensors: ``new_int(shape, fill)``,
    ``new_float(shape)`` and ``arange(n)`` create arrays of the same kind."""
    nz, ny, nx = k.shape
    n = nx * ny * nz
    idx = arange(n).reshape(nz, ny, nx)
    cols = new_int((nz, ny, nx, 7), -1)
    vals = new_float((nz, ny, nx, 7))
    diag = new_float((nz, ny, nx))
    # x faces
    cx = _harmonic(k[:, :, 1:], k[:, :, :-1])
    diag[:, :, 1:] += cx
    diag[:, :, :-1] += cx
    diag[:, :, 0] += 2.0 * k[:, :, 0]
    diag[:, :, -1] += 2.0 * k[:, :, -1]
    cols[:, :, 1:, 2] = idx[:, :, :-1]
    vals[:, :, 1:, 2] = -cx
    cols[:, :, :-1, 4] = idx[:, :, 1:]
    vals[:, :, :-1, 4] = -cx
    del cx
    # y faces
    cy = _harmonic(k[:, 1:, :], k[:, :-1, :])
    diag[:, 1:, :] += cy
    diag[:, :-1, :] += cy
    diag[:, 0, :] += 2.0 * k[:, 0, :]
    diag[:, -1, :] += 2.0 * k[:, -1, :]
    cols[:, 1:, :, 1] = idx[:, :-1, :]
    vals[:, 1:, :, 1] = -cy
    cols[:, :-1, :, 5] = idx[:, 1:, :]
    vals[:, :-1, :, 5] = -cy
    del cy
    # z faces (absent for a single plane: then the matrix is the 2D operator)
    if nz > 1:
        cz = _harmonic(k[1:, :, :], k[:-1, :, :])
        diag[1:, :, :] += cz
        diag[:-1, :, :] += cz
        cols[1:, :, :, 0] = idx[:-1, :, :]
        vals[1:, :, :, 0] = -cz
        cols[:-1, :, :, 6] = idx[1:, :, :]
        vals[:-1, :, :, 6] = -cz
        del cz
        diag[0, :, :] += 2.0 * k[0, :, :]
        diag[-1, :, :] += 2.0 * k[-1, :, :]
    cols[..., 3] = idx
    vals[..., 3] = diag
    return cols.reshape(n, 7), vals.reshape(n, 7)


def diffusion_3d(nx: int, ny: int | None = None, nz: int | None = None, sigma: float = 1.0,
                 seed: int = 1) -> CsrMatrix:
    """7-point variable-coefficient diffusion (SPD), sorted CSR.

    Cell (z, y, x) has row ``(z*ny + y)*nx + x``; conductivities
    ``k = exp(sigma * N(0,1))`` drawn as a (nz, ny, nx) array from
    ``default_rng(seed)``; the operator is :func:`diffusion_stencil`. Within a
    row the columns are in increasing order (-z, -y, -x, diag, +x, +y, +z).
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    rng = np.random.default_rng(seed)
    k = _lognormal_k((nz, ny, nx), sigma, rng)
    n = nx * ny * nz
    cols, vals = diffusion_stencil(k, lambda s, f: np.full(s, f, dtype=np.int64),
                                   lambda s: np.zeros(s, dtype=np.float64),
                                   lambda m: np.arange(m, dtype=np.int64))
    meta = {"kind": "diffusion_3d" if nz > 1 else "diffusion_2d", "nx": nx, "ny": ny, "nz": nz,
            "sigma": sigma, "seed": seed}
    return _stencil_to_csr(cols, vals, n, meta)


def diffusion_keys_torch(shape, sigma, gen, device="cuda"):
    """Packed keys + FP32 values of a log-normal diffusion matrix
    (:func:`diffusion_stencil`), generated with torch on ``device``.
    shape = (nz, ny, nx); nz = 1 gives the 5-point 2D operator."""
    import torch

    k = torch.exp(sigma * torch.randn(shape, generator=gen, device=device, dtype=torch.float64))
    return diffusion_keys_from_k(k)


def diffusion_keys_from_k(k):
    """(keys int64[nnz], vals float32[nnz], n) of the diffusion operator of a
    torch conductivity array k (nz, ny, nx), in row-major stencil order."""
    import torch

    dev = k.device
    cols, vals = diffusion_stencil(k, lambda s, f: torch.full(s, f, dtype=torch.int64, device=dev),
                                   lambda s: torch.zeros(s, dtype=torch.float64, device=dev),
                                   lambda m: torch.arange(m, dtype=torch.int64, device=dev))
    n = cols.shape[0]
    rows = torch.arange(n, device=dev, dtype=torch.int64)[:, None].expand(n, 7)
    m = cols >= 0
    keys = (rows[m] << 32) | cols[m]
    return keys.contiguous(), vals[m].float().contiguous(), n


def diffusion_2d(nx: int, ny: int | None = None, sigma: float = 1.5, seed: int = 0) -> CsrMatrix:
    """5-point analogue of :func:`diffusion_3d` on an (ny, nx) grid."""
    ny = nx if ny is None else ny
    return diffusion_3d(nx, ny, 1, sigma=sigma, seed=seed)


def batched_diffusion_sides(num_systems: int = 512, lo: int = 32, hi: int = 512,
                            seed: int = 0) -> np.ndarray:
    """Grid sides of the batched workload: uniform integers in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=num_systems)


def rhs(n: int, seed: int = 0, nrhs: int = 1) -> np.ndarray:
    """float32 N(0,1) right-hand side(s), shape (n,) or (nrhs, n)."""
    rng = np.random.default_rng(seed)
    if nrhs == 1:
        return rng.standard_normal(n).astype(np.float32)
    return rng.standard_normal((nrhs, n)).astype(np.float32)


def laplacian_1d_csc(n: int, shift: float = 0.1):
    """The reference's direct-solver benchmark matrix (sparse_bridge/benchmark.py:
    20-36): tridiagonal 1D Laplacian (2 on the diagonal, -1 beside it) plus
    shift * I, as CSC (col_starts int64[n+1], row_indices int64, values float64),
    rows sorted within each column."""
    cols = []
    for off in (-1, 0, 1):
        r = np.arange(n, dtype=np.int64)
        c = r + off
        keep = (c >= 0) & (c < n)
        cols.append((r[keep], c[keep], np.full(int(keep.sum()), 2.0 + shift if off == 0 else -1.0)))
    rows = np.concatenate([t[0] for t in cols])
    cc = np.concatenate([t[1] for t in cols])
    vals = np.concatenate([t[2] for t in cols])
    order = np.lexsort((rows, cc))
    rows, cc, vals = rows[order], cc[order], vals[order]
    col_starts = np.zeros(n + 1, dtype=np.int64)
    np.add.at(col_starts, cc + 1, 1)
    return np.cumsum(col_starts).astype(np.int64), rows.astype(np.int64), vals.astype(np.float64)
