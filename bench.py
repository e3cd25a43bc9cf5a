#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

Workload (the block-sparse system, n = 1,048,576; 65,536 block-rows of 16x16
blocks, diagonal block + 8 random block-columns; nnz = 150,994,944; seed 0;
synthetic, generated here):

1. COO -> CSR on the GPU from the shuffled packed u64 keys (timed, reported);
2. CSR SpMV y = A x, FP32 -- the timed steps: `--steps` launches after
   `--warmup`, bracketed by HIP events on the launch stream. `value` is the
   SpMV's HBM throughput from algorithmic bytes nnz*(4+4) + 4(n+1) + 4n + 4n;
3. BiCGSTAB to relative residual 1e-6 (time-to-solution; true residual
   checked in FP64 on the GPU), and CG on the 64x64 Laplacian;
4. CPU baseline: scipy CSR matvec / BiCGSTAB on the host cores (oracle leg).

Prints one JSON line. `--gpus` above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

METRIC = "time-to-solution at relres 1e-6: learned-PCG vs plain CG; SpMV HBM GB/s"
PEAK_HBM_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def spmv_bytes(n, m, nnz, val_bytes=4):
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 4 * m + 4 * n


def measured_traffic(key):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(HERE, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(key, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def time_launches(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def true_relres_gpu(rows, cols, vals, x, b):
    """||b - A x|| / ||b|| in FP64 with torch index_add (independent of the library)."""
    import torch

    ax = torch.zeros(b.numel(), dtype=torch.float64, device=b.device)
    ax.index_add_(0, rows, vals.double() * x.double()[cols])
    r = b.double() - ax
    return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b.double()))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sections", default="spmv,solve,cpu",
                    help="comma list of spmv, solve, cpu, copy (copy: 1 GiB device copy for counter calibration)")
    ap.add_argument("--block-rows", type=int, default=65536, help="block-rows of the block-sparse system")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    sections = set(args.sections.split(","))

    import numpy as np
    import torch

    from mpsparse_amd import problems, spmv

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    out = {"metric": METRIC, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded generator, random-init values)"}

    if "copy" in sections:
        nbytes = 1 << 30
        src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        ms = time_launches(lambda: dst.copy_(src), 10, 2)
        out["copy_calibration"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
                                   "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
        log("copy", out["copy_calibration"])

    t0 = time.time()
    q = problems.block_sparse(args.block_rows, 16, 8, seed=0)
    n, nnz = q.num_rows, q.nnz
    log(f"generated n={n} nnz={nnz} in {time.time() - t0:.1f}s")
    keys = torch.from_numpy(q.keys.view(np.int64)).to(dev)
    vals = torch.from_numpy(q.vals).to(dev)
    b = torch.from_numpy(q.b).to(dev)
    del q
    empty = torch.empty(0)
    # COO -> CSR (first call includes one-time allocation; time a second build)
    A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
    e1.record()
    torch.cuda.synchronize()
    coo_ms = e0.elapsed_time(e1)
    log(f"coo->csr {coo_ms:.2f} ms")
    rp = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ci = torch.empty(nnz, dtype=torch.int32, device=dev)
    cv = torch.empty(nnz, dtype=torch.float32, device=dev)
    A.export(rp, ci, cv)
    del keys, vals
    rows = torch.repeat_interleave(torch.arange(n, device=dev), (rp[1:] - rp[:-1]).long())
    cols = ci.long()
    config = {"workload": "block_sparse_bicgstab_csr", "n": n, "nnz": nnz, "block": 16, "blocks_per_row": 9,
              "seed": 0}
    out["config"] = config
    out["coo_to_csr"] = {"ms": coo_ms, "keys_per_s": nnz / (coo_ms * 1e-3),
                         "algorithmic_GBps": (20 * nnz + 4 * (n + 1)) / (coo_ms * 1e-3) / 1e9}

    if "spmv" in sections:
        x = torch.randn(n, device=dev)
        y = torch.empty(n, device=dev)
        ms = time_launches(lambda: A.mv(x, y), args.steps, args.warmup)
        alg = spmv_bytes(n, n, nnz)
        gbs = alg / (ms * 1e-3) / 1e9
        # spot check of this y against an FP64 product computed by torch
        y64 = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(0, rows, cv.double() * x.double()[cols])
        err = float((y.double() - y64).abs().max() / y64.abs().max())
        out["value"] = gbs
        out["ms_per_step"] = ms
        traffic = measured_traffic("csr_spmv_block_sparse")
        out["roofline"] = {"bound": "hbm", "kernel": "k_spmv<ValF32,0,XPlain,EpiStore>", "achieved": gbs,
                           "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                           "traffic": traffic, "algorithmic_bytes": alg}
        out["spmv_check"] = {"max_rel_err_vs_fp64": err}
        log(f"spmv {ms:.4f} ms  {gbs:.1f} GB/s  err {err:.2e}")

    if "solve" in sections:
        x = torch.zeros(n, device=dev)
        A.bicgstab(b, x)  # warm-up (workspace allocation)
        times, info = [], None
        for _ in range(3):
            x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            info = A.bicgstab(b, x)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        rel = true_relres_gpu(rows, cols, cv, x, b)
        out["time_to_solution"] = {"solver": "bicgstab", "tol": 1e-6, "wall_s": float(np.median(times)),
                                   "device_s": info["seconds"], "iterations": info["iterations"],
                                   "status": info["status"], "true_relres_fp64": rel}
        log("bicgstab", out["time_to_solution"])
        # CG on the 64x64 Laplacian (latency-bound case)
        p = problems.laplacian_2d(64)
        L = spmv.csr_tensor(torch.from_numpy(p.keys.view(np.int64)).to(dev), torch.from_numpy(p.vals).to(dev),
                            empty, empty, empty, p.num_rows, p.num_cols)
        bl = torch.from_numpy(p.b).to(dev)
        xl = torch.zeros(p.num_rows, device=dev)
        L.cgd(bl, xl)
        ct = []
        for _ in range(5):
            torch.cuda.synchronize()
            t = time.perf_counter()
            cinfo = L.cgd(bl, xl)
            torch.cuda.synchronize()
            ct.append(time.perf_counter() - t)
        out["cg_laplacian_64"] = {"wall_s": float(np.median(ct)), "iterations": cinfo["iterations"],
                                  "status": cinfo["status"], "relres": cinfo["relres"]}
        log("cg laplacian", out["cg_laplacian_64"])

    if "cpu" in sections:
        import scipy.sparse as sp

        from oracle import krylov as K

        cores = len(os.sched_getaffinity(0))
        S = sp.csr_matrix((cv.cpu().numpy(), ci.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
        xh = np.random.default_rng(0).standard_normal(n).astype(np.float32)
        S @ xh
        reps = 5
        t = time.perf_counter()
        for _ in range(reps):
            S @ xh
        cpu_ms = (time.perf_counter() - t) / reps * 1e3
        out["cpu_baseline"] = {"value": spmv_bytes(n, n, nnz) / (cpu_ms * 1e-3) / 1e9, "unit": "GB/s",
                               "cores": 1, "kind": "port",
                               "sample": f"scipy CSR matvec (single-threaded) on the full matrix, {reps} calls, "
                                         f"{cpu_ms:.1f} ms each; host has {cores} cores"}
        log("cpu spmv", out["cpu_baseline"])
        if "solve" in sections:
            t = time.perf_counter()
            xc, itc, infoc = K.scipy_bicgstab(S, b.cpu().numpy(), tol=1e-6)
            tc = time.perf_counter() - t
            out["cpu_baseline"]["bicgstab"] = {"wall_s": tc, "iterations": itc, "info": int(infoc),
                                               "cores_available": cores}
            log("cpu bicgstab", tc, itc)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
