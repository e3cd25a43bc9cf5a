"""Generators of the benchmark systems, checked against independent constructions (CPU)."""
import numpy as np
import scipy.sparse as sp

from mpsparse_amd import problems as P
from oracle import coo_csr as oc


def test_laplacian_matches_kron_construction():
    g = 64
    L1 = sp.diags([-np.ones(g - 1), 2 * np.ones(g), -np.ones(g - 1)], [-1, 0, 1])
    A_ref = (sp.kron(sp.eye(g), L1) + sp.kron(L1, sp.eye(g))).tocsr()
    A_ref.sort_indices()
    p = P.laplacian_2d(g)
    assert p.nnz == 5 * g * g - 4 * g == 20224
    rp, ci, v = oc.coo_to_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    assert np.array_equal(rp.astype(np.int64), A_ref.indptr)
    assert np.array_equal(ci.astype(np.int64), A_ref.indices)
    assert np.array_equal(v.astype(np.float64), A_ref.data)
    # the keys really are shuffled and b is N(0,1)-ish
    assert not np.all(np.diff(p.keys.astype(np.float64)) >= 0)
    assert abs(float(p.b.mean())) < 0.1 and abs(float(p.b.std()) - 1) < 0.1


def test_diffusion_3d_structure():
    nx = 6
    A = P.diffusion_3d(nx, sigma=1.0, seed=1)
    n = nx ** 3
    assert A.nnz == 7 * n - 6 * nx * nx
    S = A.to_scipy()
    assert abs(S - S.T).max() == 0
    assert np.linalg.eigvalsh(S.toarray()).min() > 0
    # interior rows sum to zero; boundary rows carry the Dirichlet faces
    rs = np.asarray(S.sum(axis=1)).ravel().reshape(nx, nx, nx)
    assert np.allclose(rs[1:-1, 1:-1, 1:-1], 0, atol=1e-5)
    assert np.all(rs[0] > 0) and np.all(rs[:, :, -1] > 0)


def test_diffusion_3d_full_size_counts():
    # only the arithmetic: n and nnz of the 256^3 system
    n = 256 ** 3
    assert 7 * n - 6 * 256 * 256 == 117047296


def test_block_sparse_structure():
    q = P.block_sparse(32, 16, 8, seed=0)
    assert q.nnz == 32 * 9 * 256
    rows = (q.keys >> np.uint64(32)).astype(np.int64)
    cols = (q.keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    A = sp.coo_matrix((q.vals.astype(np.float64), (rows, cols)), shape=(q.num_rows,) * 2).tocsr()
    assert A.nnz == q.nnz  # no duplicate keys
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    assert np.allclose(d - off, 1.0, atol=1e-4)
    # every block row touches 9 distinct block columns including its own
    br = rows // 16
    bc = cols // 16
    for r in range(32):
        s = np.unique(bc[br == r])
        assert s.size == 9 and r in s
    assert abs(A - A.T).max() > 0  # nonsymmetric


def test_full_size_block_sparse_counts():
    assert 65536 * 9 * 256 == 150994944


def test_torch_diffusion_generator_matches_numpy():
    """The torch (device) generator used for training builds the same matrix as
    diffusion_3d from the same conductivities (2D and 3D)."""
    import torch

    from mpsparse_amd.problems import diffusion_keys_from_k

    for shape in ((4, 5, 6), (1, 7, 9)):
        nz, ny, nx = shape
        A = P.diffusion_3d(nx, ny, nz, sigma=1.3, seed=3)
        k = P._lognormal_k(shape, 1.3, np.random.default_rng(3))
        keys, vals, n = diffusion_keys_from_k(torch.from_numpy(k))
        rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(A.row_ptr.astype(np.int64)))
        assert n == A.num_rows
        assert np.array_equal(keys.numpy().view(np.uint64), P.pack_keys(rows, A.col_ind.astype(np.uint64)))
        assert np.array_equal(vals.numpy(), A.vals)
