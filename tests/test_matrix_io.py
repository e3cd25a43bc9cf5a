"""Matrix Market and npz readers/writers against scipy's (CPU)."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from mpsparse_amd import matrix_io as MIO
from mpsparse_amd.problems import unpack_keys
from oracle import coo_csr as oc


def _dense(keys, vals, n, m):
    r, c = unpack_keys(keys)
    d = np.zeros((n, m))
    np.add.at(d, (r.astype(np.int64), c.astype(np.int64)), vals.astype(np.float64))
    return d


def _rand(n, m, density, seed):
    return sp.random(n, m, density=density, random_state=np.random.default_rng(seed), dtype=np.float64,
                     data_rvs=lambda k: np.random.default_rng(seed + 1).standard_normal(k)).astype(np.float32)


def test_read_general_and_symmetric(tmp_path):
    A = _rand(40, 55, 0.1, 1)
    scipy.io.mmwrite(tmp_path / "g.mtx", A)
    keys, vals, n, m = MIO.read_matrix_market(tmp_path / "g.mtx")
    assert (n, m) == (40, 55) and keys.size == A.nnz
    assert np.allclose(_dense(keys, vals, n, m), A.toarray(), rtol=1e-7, atol=0)
    S = _rand(30, 30, 0.15, 2)
    S = (S + S.T).tocoo()
    scipy.io.mmwrite(tmp_path / "s.mtx", S, symmetry="symmetric")
    keys, vals, n, m = MIO.read_matrix_market(tmp_path / "s.mtx")
    assert np.allclose(_dense(keys, vals, n, m), S.toarray(), rtol=1e-6, atol=0)


def test_pattern_skew_and_errors(tmp_path):
    (tmp_path / "p.mtx").write_text("%%MatrixMarket matrix coordinate pattern general\n% c\n3 4 3\n1 1\n2 4\n3 2\n")
    keys, vals, n, m = MIO.read_matrix_market(tmp_path / "p.mtx")
    assert (n, m) == (3, 4) and np.all(vals == 1)
    r, c = unpack_keys(keys)
    assert r.tolist() == [0, 1, 2] and c.tolist() == [0, 3, 1]
    (tmp_path / "k.mtx").write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2\n")
    d = _dense(*MIO.read_matrix_market(tmp_path / "k.mtx"))
    assert np.allclose(d, -d.T) and d[1, 0] == 1.5 and d[0, 1] == -1.5
    (tmp_path / "a.mtx").write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ValueError, match="coordinate"):
        MIO.read_matrix_market(tmp_path / "a.mtx")
    (tmp_path / "o.mtx").write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(ValueError, match="outside"):
        MIO.read_matrix_market(tmp_path / "o.mtx")


def test_write_round_trips_through_scipy(tmp_path):
    A = _rand(64, 48, 0.08, 3).tocsr()
    A.sort_indices()
    MIO.write_matrix_market(tmp_path / "w.mtx.gz", A.indptr, A.indices, A.data, 64, 48, comment="test")
    B = scipy.io.mmread(tmp_path / "w.mtx.gz")
    assert np.array_equal(B.toarray().astype(np.float32), A.toarray())
    keys, vals, n, m = MIO.read_matrix_market(tmp_path / "w.mtx.gz")
    rp, ci, v = oc.coo_to_csr(keys, vals, n, m)
    assert np.array_equal(rp.astype(np.int64), A.indptr) and np.array_equal(ci.astype(np.int64), A.indices)
    assert np.array_equal(v, A.data)  # FP32 values survive the text round trip exactly


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo"])
def test_npz_both_ways(tmp_path, fmt):
    A = _rand(50, 37, 0.1, 4).asformat(fmt)
    sp.save_npz(tmp_path / "a.npz", A, compressed=False)
    keys, vals, n, m = MIO.load_npz(tmp_path / "a.npz")
    assert (n, m) == (50, 37) and np.array_equal(_dense(keys, vals, n, m), A.toarray().astype(np.float64))
    C = A.tocsr()
    MIO.save_npz(tmp_path / "b.npz", C.indptr, C.indices, C.data, 50, 37)
    assert np.array_equal(sp.load_npz(tmp_path / "b.npz").toarray(), C.toarray())
