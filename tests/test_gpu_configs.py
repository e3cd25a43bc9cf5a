"""The five benchmark workloads at their full stated sizes, on the GPU, each
checked against the oracle or an FP64 recomputation independent of the library.

* Block-sparse BiCGSTAB, n = 1,048,576 (65,536 block-rows of 16x16 blocks, the
  diagonal block plus 8 random block-columns; 150,994,944 keys, seed 0):
  COO -> CSR from the shuffled packed keys bit-exact against the stable-sort
  oracle (oracle.coo_to_csr_bucketed, pinned to numpy's stable argsort);
  CSR and BCSR-16 y = A x within 2 gamma_k (|A||x|) of the FP64 product
  (k = 144 entries per row); BiCGSTAB (CSR and BCSR-16) to FP64 true relres
  <= 1.05e-6 with the iteration count within 15% + 2 of the FP64 oracle's.
* 3D 7-point log-normal diffusion 256^3 (sigma 1, seed 1; n = 16,777,216):
  FP64 iterative refinement with FP32, BF16 and FP16 stored values to FP64
  true relres <= 1.05e-10.
* Batched Jacobi-PCG: 512 2D log-normal diffusion systems (sides uniform in
  [32, 512], sigma 1.5, seed 0), 4 right-hand sides each: every one of the
  2048 pairs to FP64 true relres <= 1.05e-6 (recomputed in FP64 with torch).
* Learned FSAI with the shipped stated-5 weights on the held-out 128^3
  sigma = 2 system (torch seed 2024): no non-positive diagonal of G; one FP32
  FSAI-PCG recurrence to FP64 true relres <= 1.05e-6 with its step count within
  15% + 2 of FP64 PCG with the same G; FP64 refinement (FP32 and BF16 values)
  to FP64 true relres <= 1.05e-6 in at most 4 outer steps and 1.35x the FP64
  step count in all (every outer step restarts the recurrence).
* A matrix written as Matrix Market and as .npz, read back by
  mpsparse_amd.matrix_io and built on the GPU: COO -> CSR bit-exact, CG to
  FP64 true relres <= 1.05 tol.

FP64 products here use scipy on the host or torch index_add on the device;
neither touches the library.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from gpu_util import say
from oracle import coo_csr as oc
from oracle import krylov as K
from oracle import spmv as OS

pytestmark = pytest.mark.gpu

WORKERS = min(16, len(os.sched_getaffinity(0)))


def _torch():
    import torch

    return torch


# ---------------------------------------------------------------------------
# block-sparse, n = 1,048,576
@pytest.fixture(scope="module")
def block_sparse_1m():
    torch = _torch()
    from mpsparse_amd import problems, spmv

    say("block-sparse 1M: generate")
    q = problems.block_sparse(65536, 16, 8, seed=0)
    n, nnz = q.num_rows, q.nnz
    assert nnz == 150_994_944
    say("block-sparse 1M: oracle stable sort")
    rp_ref, ci_ref, v_ref = oc.coo_to_csr_bucketed(q.keys, q.vals, n, n, workers=WORKERS)
    say("block-sparse 1M: GPU build")
    rp = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    ci = torch.zeros(nnz, dtype=torch.int32, device="cuda")
    cv = torch.zeros(nnz, dtype=torch.float32, device="cuda")
    A = spmv.csr_tensor(torch.from_numpy(q.keys.view(np.int64)).cuda(), torch.from_numpy(q.vals).cuda(), rp, ci,
                        cv, n, n)
    S = sp.csr_matrix((v_ref.astype(np.float64), ci_ref.astype(np.int64), rp_ref.astype(np.int64)), shape=(n, n))
    d = {"q": q, "n": n, "nnz": nnz, "ref": (rp_ref, ci_ref, v_ref), "A": A, "S": S,
         "got": (rp.cpu().numpy().view(np.uint32), ci.cpu().numpy().view(np.uint32), cv.cpu().numpy())}
    del rp, ci, cv
    yield d


def test_block_sparse_1m_coo_to_csr_bit_exact(block_sparse_1m):
    d = block_sparse_1m
    for name, got, ref in zip(("row_ptr", "col_ind", "values"), d["got"], d["ref"]):
        assert got.shape == ref.shape, name
        bad = np.flatnonzero(got != ref)
        assert bad.size == 0, (name, bad[:5])


def _spmv_check(S, row_ptr, col_ind, vals, x, y):
    yref = S @ x.astype(np.float64)
    bound = OS.spmv_error_bound(row_ptr, col_ind, vals, x)
    err = np.abs(y.astype(np.float64) - yref)
    assert np.all(err <= 2 * bound + 1e-30), (float(err.max()), float(bound[np.argmax(err - 2 * bound)]))


def test_block_sparse_1m_spmv_csr_and_bcsr16(block_sparse_1m):
    torch = _torch()
    from mpsparse_amd import bcsr

    d = block_sparse_1m
    n = d["n"]
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(n, device="cuda")
    say("block-sparse 1M: CSR mv")
    d["A"].mv(xt, y)
    rp, ci, v = d["ref"]
    _spmv_check(d["S"], rp, ci, v, x, y.cpu().numpy())
    say("block-sparse 1M: BCSR-16 build + mv")
    q = d["q"]
    Bm = bcsr.bcsr_tensor(torch.from_numpy(q.keys.view(np.int64)).cuda(), torch.from_numpy(q.vals).cuda(),
                          torch.empty(0, dtype=torch.int32), torch.empty(0, dtype=torch.int32), n, n, 16)
    assert Bm.num_blocks == 589_824
    yb = torch.empty(n, device="cuda")
    Bm.mv(xt, yb)
    _spmv_check(d["S"], rp, ci, v, x, yb.cpu().numpy())
    d["B"] = Bm


@pytest.mark.parametrize("fmt", ["csr", "bcsr16"])
def test_block_sparse_1m_bicgstab(block_sparse_1m, fmt):
    torch = _torch()
    d = block_sparse_1m
    q, S = d["q"], d["S"]
    M = d["A"] if fmt == "csr" else d.get("B")
    if M is None:
        pytest.skip("BCSR-16 handle is built by the SpMV test")
    say("block-sparse 1M: bicgstab", fmt)
    x = torch.full((d["n"],), float("nan"), device="cuda")
    info = M.bicgstab(torch.from_numpy(q.b).cuda(), x)
    assert info["status"] == "converged", info
    xs = x.cpu().numpy()
    rel = K.true_relres(S, xs, q.b)
    assert rel <= 1.05e-6, (rel, info)
    if "it64" not in d:
        _, d["it64"], _, _ = K.bicgstab(S, q.b.astype(np.float64), tol=1e-6)
    it64 = d["it64"]
    assert abs(info["iterations"] - it64) <= 0.15 * it64 + 2, (info, it64)
    print(f"[block-sparse 1M] {fmt} bicgstab {info['iterations']} its (FP64 oracle {it64}), relres {rel:.2e}, "
          f"device {info['seconds'] * 1e3:.2f} ms")


# ---------------------------------------------------------------------------
# 3D diffusion 256^3, FP64 refinement to 1e-10
@pytest.fixture(scope="module")
def diffusion_256():
    torch = _torch()
    from mpsparse_amd import problems, spmv

    say("diffusion 256^3: generate")
    M = problems.diffusion_3d(256, sigma=1.0, seed=1)
    n = M.num_rows
    rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(M.row_ptr.astype(np.int64)))
    keys = problems.pack_keys(rows, M.col_ind)
    del rows
    say("diffusion 256^3: GPU build")
    e = torch.empty(0)
    A = spmv.csr_tensor(torch.from_numpy(keys.view(np.int64)).cuda(), torch.from_numpy(M.vals).cuda(), e, e, e, n, n)
    del keys
    yield {"M": M, "S": M.to_scipy(), "A": A, "b": problems.rhs(n, seed=1).astype(np.float64)}


@pytest.mark.parametrize("values", ["f32", "bf16", "f16"])
def test_diffusion_256_refined_to_1e_10(diffusion_256, values):
    torch = _torch()
    d = diffusion_256
    b = d["b"]
    x = torch.zeros(b.size, dtype=torch.float64, device="cuda")
    say("diffusion 256^3: refined", values)
    info = d["A"].solve_refined(torch.from_numpy(b).cuda(), x, tol=1e-10, values=values)
    assert info["status"] == "converged", info
    rel = K.true_relres(d["S"], x.cpu().numpy(), b)
    assert rel <= 1.05e-10, (rel, info)
    print(f"[diffusion 256^3] {values}: {info['outer_iterations']} outer / {info['inner_iterations']} inner steps, "
          f"relres {rel:.2e}")


@pytest.mark.parametrize("values", ["f32", "bf16", "f16"])
def test_diffusion_256_marched_product_bit_identical(diffusion_256, values):
    """At 256^3 the symmetric storage's product walks the rows along the
    +-65536 diagonal pair (the far pair's values and x carried in registers
    from one grid plane to the next). Against the same matrix with one lower
    entry of its last row changed (no longer symmetric: 7 planes, the
    row-group walk) every other row's product agrees bit for bit; the FP32
    product stays within 2 gamma_k (|A||x|) of the FP64 product."""
    torch = _torch()
    from gpu_util import csr_from_arrays

    M = diffusion_256["M"]
    n = M.num_rows
    if "perturbed" not in diffusion_256:
        v2 = M.vals.copy()
        v2[M.row_ptr[n - 1]] *= 1.5  # first entry of the last row: a lower-triangle coupling
        say("diffusion 256^3: perturbed copy")
        diffusion_256["perturbed"] = csr_from_arrays(M.row_ptr, M.col_ind, v2, n, M.num_cols)
        diffusion_256["x"] = np.random.default_rng(11).standard_normal(n).astype(np.float32)
    A, B, x = diffusion_256["A"], diffusion_256["perturbed"], diffusion_256["x"]
    xt = torch.from_numpy(x).cuda()
    ya = torch.full((n,), float("nan"), device="cuda")
    yb = torch.full((n,), float("nan"), device="cuda")
    say("diffusion 256^3: products", values)
    A.mv(xt, ya, values=values)
    B.mv(xt, yb, values=values)
    assert A.format == "dia_symmetric" and B.format == "dia"
    ya, yb = ya.cpu().numpy(), yb.cpu().numpy()
    assert np.array_equal(ya[:-1].view(np.uint32), yb[:-1].view(np.uint32))
    if values == "f32":
        y64 = OS.csr_spmv(M.row_ptr, M.col_ind, M.vals, x)
        bound = 2 * OS.spmv_error_bound(M.row_ptr, M.col_ind, M.vals, x) + 1e-30
        assert np.all(np.abs(ya.astype(np.float64) - y64) <= bound)


# ---------------------------------------------------------------------------
# batched Jacobi-PCG, 512 systems x 4 right-hand sides
def test_batched_512_systems_all_pairs():
    torch = _torch()
    from mpsparse_amd import batched, spmv

    say("batched 512: generate")
    keys, vals, off, sides = batched.batched_diffusion(512, 32, 512, 1.5, 0, "cuda")
    n, nnz = int(off[-1]), keys.numel()
    e = torch.empty(0)
    A = spmv.csr_tensor(keys, vals, e, e, e, n, n)
    rows, cols = (keys >> 32).long(), (keys & 0xFFFFFFFF).long()
    del keys
    B = batched.batch_tensor(A, torch.from_numpy(off))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1)
    b = torch.randn((n, 4), generator=gen, device="cuda")
    x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    say("batched 512: solve", n, nnz)
    info = B.pcg(b, x)
    assert info["converged_pairs"] == info["total_pairs"] == 2048, info
    assert info["diagonal_planes"] == 5 and info["symmetric_storage"] == 1, info
    sys_of_row = torch.repeat_interleave(torch.arange(512, device="cuda"), torch.from_numpy(np.diff(off)).cuda())
    ax = torch.zeros((n, 4), dtype=torch.float64, device="cuda").index_add_(0, rows,
                                                                           vals.double()[:, None] * x[cols])
    r = b.double() - ax
    rr = torch.zeros((512, 4), dtype=torch.float64, device="cuda").index_add_(0, sys_of_row, r * r)
    bb = torch.zeros((512, 4), dtype=torch.float64, device="cuda").index_add_(0, sys_of_row, b.double() ** 2)
    rel = (rr / bb).sqrt().cpu().numpy()
    assert rel.max() <= 1.05e-6, (rel.max(), np.unravel_index(rel.argmax(), rel.shape))
    reported = info["relres"].numpy()
    assert np.all(np.abs(reported - rel) <= 0.01 * rel + 1e-12)
    print(f"[batched 512] n={n} nnz={nnz}: max relres {rel.max():.3e}, max its {info['max_iterations']}, "
          f"device {info['seconds']:.2f} s")


# ---------------------------------------------------------------------------
# learned FSAI, held-out 128^3 sigma 2, shipped stated-5 weights
def _fp64_pcg_same_g(rows, cols, vals, grows, gcols, g, b, tol, maxiter=5000):
    """FP64 PCG with M = G^T G in torch (index_add products); returns iterations."""
    torch = _torch()
    n = b.numel()

    def A(v):
        return torch.zeros(n, dtype=torch.float64, device=v.device).index_add_(0, rows, vals * v[cols])

    def Mv(v):
        w = torch.zeros(n, dtype=torch.float64, device=v.device).index_add_(0, grows, g * v[gcols])
        return torch.zeros(n, dtype=torch.float64, device=v.device).index_add_(0, gcols, g * w[grows])

    x = torch.zeros(n, dtype=torch.float64, device=b.device)
    r = b.clone()
    z = Mv(r)
    p = z.clone()
    rz = torch.dot(r, z)
    thr = tol * float(torch.linalg.vector_norm(b))
    for it in range(1, maxiter + 1):
        q = A(p)
        a = rz / torch.dot(p, q)
        x += a * p
        r -= a * q
        if float(torch.linalg.vector_norm(r)) <= thr:
            return it
        z = Mv(r)
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return maxiter


def test_learned_fsai_128_heldout_shipped_weights():
    torch = _torch()
    from mpsparse_amd import learned_fsai as LF
    from mpsparse_amd import spmv

    gen = torch.Generator(device="cuda")
    gen.manual_seed(2024)
    say("learned fsai 128^3: generate")
    keys, vals, n = LF.diffusion_keys_torch((128, 128, 128), 2.0, gen, "cuda")
    e = torch.empty(0)
    A = spmv.csr_tensor(keys, vals, e, e, e, n, n)
    rows, cols, v64 = (keys >> 32).long(), (keys & 0xFFFFFFFF).long(), vals.double()
    del keys
    b = torch.randn(n, generator=gen, dtype=torch.float64, device="cuda")
    net = LF.load_network()
    assert net.feature_set == "stated5"
    say("learned fsai 128^3: build G")
    F, nonpos = LF.build(A, net)
    assert nonpos == 0
    gr, gc = (t.long() for t in F.edges())
    g = F.values().double()
    say("learned fsai 128^3: FP64 PCG with the same G")
    it64 = _fp64_pcg_same_g(rows, cols, v64, gr, gc, g, b, 1e-6)
    # one FP32 recurrence (no refinement) against FP64 PCG with the same G
    x32 = torch.full((n,), float("nan"), device="cuda")
    say("learned fsai 128^3: FP32 PCG")
    info = LF.solve(A, F, b.float(), x32, tol=1e-6)
    assert info["status"] == "converged" and not info["fell_back_to_cg"], info
    ax = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, rows, v64 * x32.double()[cols])
    rel = float(torch.linalg.vector_norm(b.float().double() - ax) / torch.linalg.vector_norm(b.float().double()))
    assert rel <= 1.05e-6, (rel, info)
    assert abs(info["iterations"] - it64) <= 0.15 * it64 + 2, (info["iterations"], it64)
    print(f"[learned fsai 128^3] FP32 PCG {info['iterations']} steps (FP64 PCG same G: {it64}), relres {rel:.2e}")
    for values in ("f32", "bf16"):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        say("learned fsai 128^3: refined", values)
        info = LF.solve_refined(A, F, b, x, tol=1e-6, values=values)
        assert info["status"] == "converged" and info["fallbacks"] == 0, info
        ax = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, rows, v64 * x[cols])
        rel = float(torch.linalg.vector_norm(b - ax) / torch.linalg.vector_norm(b))
        assert rel <= 1.05e-6, (values, rel, info)
        # each outer step restarts the Krylov recurrence from the FP64 residual
        inner, outer = info["inner_iterations"], info["outer_iterations"]
        assert inner <= 1.35 * it64 and outer <= 4, (values, inner, outer, it64)
        print(f"[learned fsai 128^3] {values}: {outer} outer / {inner} inner steps (FP64 PCG same G: {it64}), "
              f"relres {rel:.2e}")


# ---------------------------------------------------------------------------
# matrices from files, through the GPU build
@pytest.mark.parametrize("fmt", ["mtx", "mtx.gz", "npz"])
def test_matrix_file_through_gpu(tmp_path, fmt):
    import scipy.io

    torch = _torch()
    from mpsparse_amd import matrix_io, problems, spmv

    M = problems.diffusion_2d(60, sigma=1.5, seed=8)
    S = M.to_scipy().astype(np.float32)
    path = str(tmp_path / f"a.{fmt}")
    if fmt == "npz":
        sp.save_npz(path, S.tocsr())
        keys, vals, n, m = matrix_io.load_npz(path)
    else:
        scipy.io.mmwrite(str(tmp_path / "a.mtx"), S.tocoo(), symmetry="symmetric")
        if fmt == "mtx.gz":
            import gzip
            import shutil

            with open(str(tmp_path / "a.mtx"), "rb") as fi, gzip.open(path, "wb") as fo:
                shutil.copyfileobj(fi, fo)
        keys, vals, n, m = matrix_io.read_matrix_market(path)
    assert (n, m) == S.shape
    perm = np.random.default_rng(9).permutation(keys.size)  # any input order
    keys, vals = keys[perm], vals[perm]
    rp_ref, ci_ref, v_ref = oc.coo_to_csr(keys, vals, n, m)
    rp = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    ci = torch.zeros(keys.size, dtype=torch.int32, device="cuda")
    cv = torch.zeros(keys.size, dtype=torch.float32, device="cuda")
    say("matrix file", fmt)
    A = spmv.csr_tensor(*matrix_io.to_csr_tensor_args((keys, vals, n, m))[:2], rp, ci, cv, n, m)
    assert np.array_equal(rp.cpu().numpy().view(np.uint32), rp_ref)
    assert np.array_equal(ci.cpu().numpy().view(np.uint32), ci_ref)
    assert np.array_equal(cv.cpu().numpy(), v_ref)
    b = np.random.default_rng(10).standard_normal(n).astype(np.float32)
    x = torch.zeros(n, device="cuda")
    info = A.cgd(torch.from_numpy(b).cuda(), x, tol=1e-5)
    Sref = sp.csr_matrix((v_ref.astype(np.float64), ci_ref.astype(np.int64), rp_ref.astype(np.int64)), shape=(n, m))
    assert info["status"] == "converged" and K.true_relres(Sref, x.cpu().numpy(), b) <= 1.05e-5, info
