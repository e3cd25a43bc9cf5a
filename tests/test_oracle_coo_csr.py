"""Pins the COO->CSR oracle (CPU only).

Three independent formulations (numpy stable argsort, the C restatement of the
reference's 4-bit LSD kernels, a Python brute force) must agree bit for bit, and
all must reproduce the hand-derived known answers in tests/golden/coo_csr_kat.json
and the reference benchmark's own check against torch's COO->CSR
(spmv_kernel/coo_csr/coo_csr/coo_csr/benchmark.py:131-156). Deliberately broken
variants are shown to fail the known answers, so the fixtures do pin the
properties they are meant to (stability, row/col roles, row_ptr offsets).
"""
import numpy as np
import pytest

from mpsparse_amd.problems import laplacian_2d, pack_keys
from oracle import coo_csr as oc

IMPLS = [oc.coo_to_csr, oc.coo_to_csr_lsd, oc.coo_to_csr_bruteforce]


def kat_arrays(case):
    coo = np.array(case["coo"], dtype=np.uint64).reshape(-1, 2)
    keys = pack_keys(coo[:, 0], coo[:, 1]) if len(coo) else np.zeros(0, np.uint64)
    return keys, np.array(case["vals"], dtype=np.float32)


def check_kat(impl, case):
    keys, vals = kat_arrays(case)
    rp, ci, ov = impl(keys, vals, case["num_rows"], case["num_cols"])
    return (np.asarray(rp, np.int64).tolist() == case["row_ptr"]
            and np.asarray(ci, np.int64).tolist() == case["col_ind"]
            and np.asarray(ov, np.float64).tolist() == case["out_vals"])


@pytest.mark.parametrize("impl", IMPLS, ids=["argsort", "lsd_c", "bruteforce"])
def test_known_answers(kat_cases, impl):
    for case in kat_cases:
        assert check_kat(impl, case), case["name"]


def _random_coo(rng, n, m, nnz, dup_range=None):
    if dup_range:  # few distinct keys -> many duplicates
        rows = rng.integers(0, min(n, dup_range), nnz)
        cols = rng.integers(0, min(m, dup_range), nnz)
    else:
        rows = rng.integers(0, n, nnz)
        cols = rng.integers(0, m, nnz)
    keys = pack_keys(rows, cols)
    vals = rng.standard_normal(nnz).astype(np.float32)
    return keys, vals


@pytest.mark.parametrize("n,m,nnz,dup", [
    (1, 1, 50, None), (7, 3, 200, None), (100, 100, 3000, 8), (1000, 37, 5000, None),
    (33, 100000, 4000, None), (5000, 5000, 1000, None), (2, 17, 300, None), (70000, 3, 2000, None),
])
def test_formulations_agree(n, m, nnz, dup):
    rng = np.random.default_rng(n * 7919 + m)
    keys, vals = _random_coo(rng, n, m, nnz, dup)
    ref = oc.coo_to_csr(keys, vals, n, m)
    lsd = oc.coo_to_csr_lsd(keys, vals, n, m)
    for a, b in zip(ref, lsd):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    if nnz <= 3000:
        bf = oc.coo_to_csr_bruteforce(keys, vals, n, m)
        for a, b in zip(ref, bf):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_matches_torch_to_sparse_csr_on_unique_keys():
    """The reference's own COO->CSR check: torch.sparse_coo_tensor(...).to_sparse_csr()."""
    import torch

    p = laplacian_2d(48, seed=3)
    rows = (p.keys >> np.uint64(32)).astype(np.int64)
    cols = (p.keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    t = torch.sparse_coo_tensor(torch.from_numpy(np.stack([rows, cols])), torch.from_numpy(p.vals),
                                (p.num_rows, p.num_cols)).to_sparse_csr()
    rp, ci, ov = oc.coo_to_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    assert np.array_equal(rp.astype(np.int64), t.crow_indices().numpy())
    assert np.array_equal(ci.astype(np.int64), t.col_indices().numpy())
    assert np.array_equal(ov, t.values().numpy())


@pytest.mark.parametrize("impl", IMPLS, ids=["argsort", "lsd_c", "bruteforce"])
def test_key_range_is_rejected(impl):
    keys = pack_keys(np.array([0, 3]), np.array([0, 1]))
    with pytest.raises(oc.KeyRangeError):
        impl(keys, np.ones(2, np.float32), 3, 4)
    keys = pack_keys(np.array([0, 1]), np.array([0, 4]))
    with pytest.raises(oc.KeyRangeError):
        impl(keys, np.ones(2, np.float32), 3, 4)


# -- the known answers catch plausible mistakes --------------------------------
def _unstable(keys, vals, n, m):  # duplicates in reverse input order
    keys = np.asarray(keys)
    order = np.lexsort((-np.arange(keys.size), keys))
    sk = keys[order]
    rows = (sk >> np.uint64(32)).astype(np.int64)
    return (np.searchsorted(rows, np.arange(n + 1)).astype(np.uint32),
            (sk & np.uint64(0xFFFFFFFF)).astype(np.uint32), vals[order])


def _swapped(keys, vals, n, m):  # sorts by (col, row): row and col roles mixed up
    keys = np.asarray(keys)
    sw = ((keys & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (keys >> np.uint64(32))
    order = np.argsort(sw, kind="stable")
    sk = keys[order]
    rows = (sk >> np.uint64(32)).astype(np.int64)
    return (np.searchsorted(rows, np.arange(n + 1)).astype(np.uint32),
            (sk & np.uint64(0xFFFFFFFF)).astype(np.uint32), vals[order])


def _rowptr_right(keys, vals, n, m):  # row_ptr from searchsorted side="right"
    rp, ci, ov = oc.coo_to_csr(keys, vals, n, m)
    sk = np.sort(np.asarray(keys), kind="stable")
    rows = (sk >> np.uint64(32)).astype(np.int64)
    return np.searchsorted(rows, np.arange(n + 1), side="right").astype(np.uint32), ci, ov


@pytest.mark.parametrize("bad", [_unstable, _swapped, _rowptr_right])
def test_known_answers_reject_broken_variants(kat_cases, bad):
    assert not all(check_kat(bad, c) for c in kat_cases)


@pytest.mark.parametrize("n,m,nnz,dup", [(1, 1, 50, True), (7, 3, 200, True), (1000, 5, 30_000, True),
                                         (5000, 2 ** 32 - 1, 100_000, False), (2 ** 20, 2 ** 20, 400_000, True),
                                         (300, 300, 0, False)])
def test_bucketed_oracle_equals_stable_argsort(n, m, nnz, dup):
    """The threaded bucketed form is the same stable sort (duplicates in input
    order, empty and skewed rows, buckets wider than the row range)."""
    rng = np.random.default_rng(n + nnz)
    rows = rng.integers(0, n, nnz, dtype=np.uint64)
    if nnz:
        rows[: nnz // 3] = rows[0]  # one heavy row
    cols = rng.integers(0, min(m, 8) if dup else m, nnz, dtype=np.uint64)
    keys = (rows << np.uint64(32)) | cols
    vals = np.arange(nnz, dtype=np.float32)  # distinct values expose any reordering of duplicates
    a = oc.coo_to_csr(keys, vals, n, m)
    for workers, buckets in ((1, 2), (4, 128), (8, 256)):
        b = oc.coo_to_csr_bucketed(keys, vals, n, m, workers=workers, buckets=buckets)
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
