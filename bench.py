#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

Headline (`value`): time to FP64 true relative residual 1e-6 of learned-FSAI
PCG (the stated 5-input edge network; BF16 matrix values, FP32 Krylov, FP64
refinement; preconditioner set-up included) against plain CG on the same GPU,
as the ratio plain-CG seconds / learned-PCG seconds, on a held-out 128^3
7-point log-normal diffusion system (sigma 2, n = 2,097,152; synthetic,
generated on the GPU). A "step" is one such learned-PCG solve: `--warmup`
untimed, then `--steps` timed; ms_per_step is their median. Plain CG is the
faster of FP32 CG and FP64-refined CG (median of `--reps`). Jacobi-PCG,
classical FSAI-PCG (FP32 and BF16 values of A) and the 8-feature network
variant are reported beside it.

SpMV part (`spmv`, `roofline`): the block-sparse system, n = 1,048,576 (65,536
block-rows of 16x16 blocks, diagonal block + 8 random block-columns;
nnz = 150,994,944; seed 0): COO -> CSR on the GPU from shuffled packed u64 keys,
then `--spmv-steps` CSR SpMV launches bracketed by HIP events on the launch
stream; HBM throughput from algorithmic bytes nnz*(4+4) + 4(n+1) + 4n + 4n.
Also BiCGSTAB to 1e-6 on it, CG on the 64x64 Laplacian, BCSR-16, the 256^3
precision path (FP64 refinement to 1e-10), the 512-system batched solver and
the direct QR solver.

CPU baseline (one thread, this host): scipy Jacobi-PCG to 1e-6 on the headline
system (measured), a bounded plain scipy CG sample there, scipy CG on the 64x64
Laplacian (median of 5), scipy matvec and BiCGSTAB on the block-sparse system.

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
    """(HBM bytes per launch, where the figure comes from) from the committed
    rocprofv3 PMC summary profiles/pmc_traffic.json, or (None, None): the
    counters need their own profiler runs, so this run does not measure them."""
    path = os.path.join(HERE, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d.get(key)
        if not e:
            return None, None
        return e.get("bytes_per_launch"), "profiles/pmc_traffic.json: " + d.get("_measured", "")
    except (OSError, ValueError):
        return None, None


def streamed_planes(A, vb, n):
    """Value planes a product with the 7-point operator streams: exactly
    symmetric storage reads the diagonal and the 3 upper planes once its 7
    planes reach 16 MB (the lower diagonals read the upper planes shifted);
    otherwise all 7."""
    return 4 if A.format == "dia_symmetric" and 7 * vb * n >= (16 << 20) else 7


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
    ap.add_argument("--steps", type=int, default=10,
                    help="timed headline solves (learned-FSAI PCG, BF16 values, 128^3 sigma 2)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed headline solves before them")
    ap.add_argument("--reps", type=int, default=5, help="timed solves per solver (median) in the comparisons")
    ap.add_argument("--spmv-steps", type=int, default=50, help="timed launches of each SpMV / stencil product")
    ap.add_argument("--sections", default="fsai,spmv,solve,bcsr,cpu,precision,batched,direct",
                    help="comma list of fsai, spmv, solve, bcsr, cpu, precision, batched, direct, and for "
                         "profiling copy (1 GiB device copy, counter calibration), coo (the block-sparse build "
                         "alone), headline (the learned-FSAI BF16 solve alone), stencil (the 256^3 products "
                         "and CG steps alone)")
    ap.add_argument("--block-rows", type=int, default=65536, help="block-rows of the block-sparse system")
    ap.add_argument("--fsai-weights", default=None, help="edge-MLP weights (.npz); default: the packaged ones")
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
    fsai = None
    if "fsai" in sections:
        fsai = fsai_section(dev, args.steps, args.warmup, args.reps, args.fsai_weights)
        out["learned_pcg"] = {k: v for k, v in fsai.items() if not k.startswith("_")}
    if "headline" in sections:  # the headline solve alone (for profiling)
        out["headline_only"] = fsai_section(dev, args.steps, args.warmup, args.reps, args.fsai_weights,
                                            only_headline=True)
    if "stencil" in sections:  # the 256^3 stencil products and CG steps alone (for profiling)
        out["stencil_256"] = precision_section(dev, torch.empty(0), args.spmv_steps, only_products=True)

    if "copy" in sections:
        nbytes = 1 << 30
        src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        ms = time_launches(lambda: dst.copy_(src), 10, 2)
        out["copy_calibration"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
                                   "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
        log("copy", out["copy_calibration"])

    config = None
    cpu = None
    if sections & {"spmv", "solve", "bcsr", "cpu", "coo"}:
        t0 = time.time()
        q = problems.block_sparse(args.block_rows, 16, 8, seed=0)
        n, nnz = q.num_rows, q.nnz
        log(f"generated n={n} nnz={nnz} in {time.time() - t0:.1f}s")
        keys = torch.from_numpy(q.keys.view(np.int64)).to(dev)
        vals = torch.from_numpy(q.vals).to(dev)
        b = torch.from_numpy(q.b).to(dev)
        del q
        empty = torch.empty(0)
        # COO -> CSR (the first build includes one-time allocations): median of 5 more
        A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
        torch.cuda.synchronize()
        coo_runs = []
        for _ in range(5):
            del A
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
            e1.record()
            torch.cuda.synchronize()
            coo_runs.append(e0.elapsed_time(e1))
        coo_ms = float(np.median(coo_runs))
        log(f"coo->csr {coo_ms:.2f} ms (runs {[round(t, 3) for t in coo_runs]})")
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
        out["coo_to_csr"] = {"ms": coo_ms, "runs_ms": coo_runs, "keys_per_s": nnz / (coo_ms * 1e-3),
                             "algorithmic_bytes": 20 * nnz + 4 * (n + 1),
                             "algorithmic_GBps": (20 * nnz + 4 * (n + 1)) / (coo_ms * 1e-3) / 1e9,
                             "note": "the build also runs the matrix set-up (row blocks, longest row, diagonal-storage "
                                     "probe) and its host synchronisations"}

        if "spmv" in sections:
            x = torch.randn(n, device=dev)
            y = torch.empty(n, device=dev)
            ms = time_launches(lambda: A.mv(x, y), args.spmv_steps, 5)
            alg = spmv_bytes(n, n, nnz)
            gbs = alg / (ms * 1e-3) / 1e9
            # spot check of this y against an FP64 product computed by torch
            y64 = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(0, rows, cv.double() * x.double()[cols])
            err = float((y.double() - y64).abs().max() / y64.abs().max())
            out["spmv"] = {"GBps": gbs, "ms": ms, "workload": "block_sparse n=1048576 nnz=150994944"}
            out["value"] = gbs
            out["ms_per_step"] = ms
            traffic, tsrc = measured_traffic("csr_spmv_block_sparse")
            out["roofline"] = {"bound": "hbm", "kernel": "k_spmv<ValF32,0,XPlain,EpiStore>", "achieved": gbs,
                               "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                               "traffic": traffic, "traffic_source": tsrc, "algorithmic_bytes": alg,
                               "launches_timed": args.spmv_steps}
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

        if "bcsr" in sections:
            from mpsparse_amd import bcsr

            # the same matrix, handed over as packed keys rebuilt from the CSR
            bkeys = (rows.to(torch.int64) << 32) | cols
            Bm = bcsr.bcsr_tensor(bkeys, cv, torch.empty(0, dtype=torch.int32), torch.empty(0, dtype=torch.int32),
                                  n, n, 16)
            del bkeys
            nblk = Bm.num_blocks
            balg = nblk * (16 * 16 * 4 + 4) + 4 * (n // 16 + 1) + 8 * n
            xb = torch.randn(n, device=dev)
            yb = torch.empty(n, device=dev)
            bms = time_launches(lambda: Bm.mv(xb, yb), args.spmv_steps, 5)
            yref = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(0, rows, cv.double() * xb.double()[cols])
            berr = float((yb.double() - yref).abs().max() / yref.abs().max())
            xs = torch.zeros(n, device=dev)
            Bm.bicgstab(b, xs)
            bt = []
            for _ in range(3):
                xs.zero_()
                torch.cuda.synchronize()
                t = time.perf_counter()
                binfo = Bm.bicgstab(b, xs)
                torch.cuda.synchronize()
                bt.append(time.perf_counter() - t)
            out["bcsr16"] = {"num_blocks": nblk, "spmv_ms": bms, "algorithmic_bytes": balg,
                             "GBps": balg / (bms * 1e-3) / 1e9, "frac_of_peak": balg / (bms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                             "max_rel_err_vs_fp64": berr,
                             "bicgstab": {"wall_s": float(np.median(bt)), "iterations": binfo["iterations"],
                                          "status": binfo["status"],
                                          "true_relres_fp64": true_relres_gpu(rows, cols, cv, xs, b)}}
            log("bcsr16", out["bcsr16"])
            del Bm

        cpu = None
        if "cpu" in sections:
            cpu = cpu_block_sparse(rp, ci, cv, b, n, nnz, "solve" in sections)

    if "precision" in sections:
        out["precision_path"] = precision_section(dev, empty, args.spmv_steps)
        pm = out["precision_path"]["mv"]
        roof = {}
        for vt in ("f32", "bf16"):
            traffic, tsrc = measured_traffic("dia_%s_diffusion_256" % vt)
            roof[vt] = {"achieved": pm[vt]["GBps"], "frac": pm[vt]["GBps"] / PEAK_HBM_GBS,
                        "algorithmic_bytes": pm[vt]["algorithmic_bytes"], "ms": pm[vt]["ms"],
                        "traffic": traffic, "traffic_source": tsrc}
        out["roofline_stencil_256"] = {"bound": "hbm", "kernel": "k_dia (7 diagonals): y = A x, 256^3",
                                       "unit": "GB/s", "peak": PEAK_HBM_GBS,
                                       "bytes": "value planes the storage streams (%s) + x + y" % pm["storage"],
                                       **roof}

    if "batched" in sections:
        out["batched_pcg"] = batched_section(dev)

    if "direct" in sections:
        out["direct_qr"] = direct_section(dev)

    if fsai is not None:
        lp = fsai["learned_fsai_stated5_refined_bf16"]
        hs = fsai["headline_steps"]
        plain = min(fsai["plain_cg_fp32"]["wall_s"], fsai["plain_cg_refined"]["wall_s"])
        out["value"] = plain / hs["median_s"]
        out["unit"] = "x (plain CG seconds / learned-FSAI PCG seconds, both to FP64 true relres 1e-6)"
        out["steps"], out["warmup"] = hs["steps"], hs["warmup"]
        out["ms_per_step"] = hs["median_s"] * 1e3
        out["step"] = ("one learned-FSAI PCG solve to relres 1e-6 (stated 5-input network, BF16 matrix values, "
                       "FP32 Krylov, FP64 refinement), preconditioner set-up included; ms_per_step is the median "
                       "of the timed steps; value = median plain-CG seconds (the faster of FP32 CG and "
                       "FP64-refined CG, %d runs) / ms_per_step" % args.reps)
        out["dtype"] = "bf16 matrix values, f32 Krylov, f64 refinement"
        out["config"] = {"workload": "learned_fsai_pcg_vs_plain_cg diffusion 128^3 sigma=2",
                         "n": fsai["n"], "nnz": fsai["nnz"], "tol": 1e-6, "seed": 2024,
                         "network": "edge MLP (a_ij, a_ii, a_jj, row nnz, row |sum|) -> 64 -> 64 -> 1",
                         "spmv_workload": config}
        out["time_to_solution_s"] = {
            "learned_fsai_pcg_bf16": hs["median_s"], "plain_cg": plain,
            "jacobi_pcg": fsai["jacobi_pcg_refined"]["wall_s"], "classical_fsai_pcg": fsai["classical_fsai_refined"]["wall_s"],
            "classical_fsai_pcg_bf16": fsai["classical_fsai_refined_bf16"]["wall_s"],
            "learned_fsai_pcg_bf16_comparison_runs": lp["wall_s"]}
        out["roofline_headline"] = {"bound": "cache", "kernel": "k_dia (7 diagonals, BF16 values): q = A p",
                                    "achieved": fsai["mv_bf16"]["GBps"], "unit": "GB/s",
                                    "algorithmic_bytes": fsai["mv_bf16"]["algorithmic_bytes"],
                                    "note": fsai["mv_bf16"]["note"]}
        if cpu is not None:
            fs = cpu_fsai_system(*(t.cpu().numpy() for t in fsai["_csr"]))
            its = fsai["plain_cg_fp32"]["iterations"]
            fs["plain_cg_extrapolated_estimate"] = {
                "s": fs["plain_cg_sample"]["s_per_iteration"] * its,
                "note": f"ESTIMATE, not measured: the sample's seconds per iteration x the {its} iterations "
                        f"plain CG needs on the GPU"}
            cpu["learned_fsai_system"] = fs
    if cpu is not None:
        cpu["laplacian_64"] = cpu_laplacian_64()
        fs = cpu.get("learned_fsai_system")
        if fs is not None:
            out["cpu_baseline"] = {"value": fs["jacobi_cg"]["wall_s"],
                                   "unit": "s (scipy Jacobi-PCG time to relres 1e-6 on the 128^3 sigma 2 system)",
                                   "cores": 1, "kind": "port",
                                   "sample": "scipy.sparse.linalg.cg with M = diag(A)^-1 run to convergence on the "
                                             "headline system (FP64, one thread: scipy's CSR matvec is "
                                             "single-threaded and BLAS is limited to 1 thread); plain scipy CG "
                                             "there is a bounded sample (see details)",
                                   "details": cpu}
        else:
            out["cpu_baseline"] = {"value": cpu["spmv"]["GBps"], "unit": "GB/s", "cores": 1, "kind": "port",
                                   "sample": cpu["spmv"]["sample"], "details": cpu}
        if fsai is not None:
            fsai.pop("_csr", None)
            out["learned_pcg"] = {k: v for k, v in fsai.items() if not k.startswith("_")}

    print(json.dumps(out), flush=True)


def _cpu_threads():
    return {"cores_available": len(os.sched_getaffinity(0)), "threads_used": 1,
            "cpu": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_block_sparse(rp, ci, cv, b, n, nnz, solve):
    """scipy on the block-sparse system (one thread): CSR matvec (5 calls) and,
    with `solve`, BiCGSTAB to relres 1e-6 (one run: it takes seconds)."""
    import numpy as np
    import scipy.sparse as sp
    from threadpoolctl import threadpool_limits

    from oracle import krylov as K

    S = sp.csr_matrix((cv.cpu().numpy(), ci.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    xh = np.random.default_rng(0).standard_normal(n).astype(np.float32)
    res = {"threads": _cpu_threads()}
    with threadpool_limits(limits=1):
        S @ xh
        ts = []
        for _ in range(5):
            t = time.perf_counter()
            S @ xh
            ts.append(time.perf_counter() - t)
        ms = float(np.median(ts)) * 1e3
        res["spmv"] = {"ms": ms, "GBps": spmv_bytes(n, n, nnz) / (ms * 1e-3) / 1e9,
                       "sample": f"scipy CSR matvec (FP32, single-threaded) on the block-sparse matrix, median "
                                 f"of 5 calls"}
        if solve:
            bh = b.cpu().numpy()
            t = time.perf_counter()
            xc, itc, infoc = K.scipy_bicgstab(S, bh, tol=1e-6)
            res["bicgstab"] = {"wall_s": time.perf_counter() - t, "iterations": itc, "info": int(infoc),
                               "true_relres_fp64": K.true_relres(S, xc, bh), "runs": 1}
    log("cpu block-sparse", res)
    return res


def cpu_laplacian_64():
    """scipy CG to relres 1e-6 on the 64x64 Laplacian (one thread): 1 warm-up,
    median of 5."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from mpsparse_amd import problems
    from oracle import coo_csr as oc
    from oracle import krylov as K
    import scipy.sparse as sp

    p = problems.laplacian_2d(64)
    rp, ci, v = oc.coo_to_csr(p.keys, p.vals, p.num_rows, p.num_cols)
    S = sp.csr_matrix((v.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(p.num_rows,) * 2)
    b = p.b.astype(np.float64)
    with threadpool_limits(limits=1):
        K.scipy_cg(S, b, tol=1e-6)
        ts = []
        for _ in range(5):
            t = time.perf_counter()
            x, it, info = K.scipy_cg(S, b, tol=1e-6)
            ts.append(time.perf_counter() - t)
    res = {"wall_s": float(np.median(ts)), "runs_s": ts, "iterations": it, "info": int(info),
           "true_relres_fp64": K.true_relres(S, x, b), "solver": "scipy cg rtol=1e-6 atol=0 (FP64)"}
    log("cpu laplacian 64", res)
    return res


def cpu_fsai_system(rp, ci, cv, b32, plain_maxiter=300):
    """scipy on the headline system (one thread): Jacobi-preconditioned CG to
    relres 1e-6 (one run, it takes tens of seconds), and plain CG stopped at
    `plain_maxiter` steps (plain CG needs ~16k steps, minutes on one core); the
    extrapolation to plain CG's full step count is a separately labelled
    estimate."""
    import numpy as np
    import scipy.sparse as sp
    from threadpoolctl import threadpool_limits

    from oracle import krylov as K

    n = rp.size - 1
    S = sp.csr_matrix((cv.astype(np.float64), ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
    b = b32.astype(np.float64)
    res = {"threads": _cpu_threads()}
    with threadpool_limits(limits=1):
        t = time.perf_counter()
        x, it, info = K.scipy_cg(S, b, tol=1e-6, maxiter=20000, jacobi=True)
        res["jacobi_cg"] = {"wall_s": time.perf_counter() - t, "iterations": it, "info": int(info),
                            "true_relres_fp64": K.true_relres(S, x, b), "runs": 1,
                            "solver": "scipy cg rtol=1e-6 atol=0, M = diag(A)^-1 (FP64)"}
        log("cpu jacobi cg", res["jacobi_cg"])
        t = time.perf_counter()
        x, it, info = K.scipy_cg(S, b, tol=1e-6, maxiter=plain_maxiter)
        dt = time.perf_counter() - t
        res["plain_cg_sample"] = {"wall_s": dt, "iterations": it, "maxiter": plain_maxiter, "info": int(info),
                                  "true_relres_fp64": K.true_relres(S, x, b), "s_per_iteration": dt / max(it, 1),
                                  "solver": "scipy cg rtol=1e-6 atol=0 (FP64), stopped at maxiter"}
        log("cpu plain cg sample", res["plain_cg_sample"])
    return res


def stencil_operator64(rp, ci, v, n, nx):
    """FP64 y = A x for a 7-point grid matrix, rebuilt as 7 coefficient planes and
    shifted vector products in plain torch (the GPU FP64 oracle; no library code)."""
    import torch

    dev = v.device
    rows = torch.repeat_interleave(torch.arange(n, device=dev), (rp[1:] - rp[:-1]).long())
    off = ci.long() - rows
    shifts = [-nx * nx, -nx, -1, 0, 1, nx, nx * nx]
    coef = torch.zeros(7, n, dtype=torch.float64, device=dev)
    for k, sft in enumerate(shifts):
        m = off == sft
        coef[k].index_put_((rows[m],), v[m].double())
    assert int(sum(int((off == sft).sum()) for sft in shifts)) == ci.numel(), "not a 7-point matrix"

    def apply(x):
        y = coef[3] * x
        for k, sft in enumerate(shifts):
            if sft > 0:
                y[:-sft] += coef[k][:-sft] * x[sft:]
            elif sft < 0:
                y[-sft:] += coef[k][-sft:] * x[:sft]
        return y

    return apply


def fp64_cg_torch(apply, b, tol, maxiter, dinv=None):
    """Textbook FP64 CG (Jacobi-preconditioned when `dinv` = 1/diag is given) on
    `apply` (torch only); stops on the true-recurrence residual ||r|| <= tol ||b||.
    Returns (x, iterations)."""
    import torch

    bd = b.double()
    x = torch.zeros_like(bd)
    r = bd.clone()
    z = r * dinv if dinv is not None else r
    p = z.clone()
    rz = torch.dot(r, z)
    thr = (tol ** 2) * float(torch.dot(bd, bd))
    it = 0
    while it < maxiter:
        q = apply(p)
        a = rz / torch.dot(p, q)
        x += a * p
        r -= a * q
        it += 1
        if it % 50 == 0:
            rr = float(torch.dot(r, r))
            if rr <= thr:
                break
            if it % 1000 == 0:
                log(f"fp64 oracle cg it {it} relres {(rr / float(torch.dot(bd, bd))) ** 0.5:.2e}")
        z = r * dinv if dinv is not None else r
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it


def batched_section(dev, num_systems=512, lo=32, hi=512, sigma=1.5, seed=0):
    """512 independent 2D 5-point log-normal diffusion systems (sides uniform
    integers in [32, 512], sigma 1.5), 4 right-hand sides each, one batched
    Jacobi-PCG solve to 1e-6 per pair; every pair's true relative residual is
    recomputed in FP64 with torch (index_add over the whole batch)."""
    import numpy as np
    import torch

    from mpsparse_amd import batched, spmv

    t = time.perf_counter()
    keys, vals, off, sides = batched.batched_diffusion(num_systems, lo, hi, sigma, seed, dev)
    n, nnz = int(off[-1]), keys.numel()
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t
    empty = torch.empty(0)
    t = time.perf_counter()
    A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
    torch.cuda.synchronize()
    t_csr = time.perf_counter() - t
    rows = (keys >> 32).long()
    cols = (keys & 0xFFFFFFFF).long()
    del keys
    log(f"batched: {num_systems} systems n={n} nnz={nnz} generated in {t_gen:.1f}s, csr {t_csr:.2f}s")
    B = batched.batch_tensor(A, torch.from_numpy(off))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    b = torch.randn((n, 4), generator=gen, device=dev)
    x = torch.zeros((n, 4), dtype=torch.float64, device=dev)  # the compensated iterate's exact sum
    t = time.perf_counter()
    info = B.pcg(b, x)  # first call: tunes the tile schedule, captures the graph
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t
    times = []
    for _ in range(2):
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        info = B.pcg(b, x)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    # FP64 true residual of every (system, rhs) pair
    sys_of_row = torch.repeat_interleave(torch.arange(num_systems, device=dev),
                                         torch.from_numpy(np.diff(off)).to(dev))
    ax = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    ax.index_add_(0, rows, vals.double()[:, None] * x.double()[cols])
    r = b.double() - ax
    rr = torch.zeros((num_systems, 4), dtype=torch.float64, device=dev).index_add_(0, sys_of_row, r * r)
    bb = torch.zeros((num_systems, 4), dtype=torch.float64, device=dev).index_add_(0, sys_of_row,
                                                                                    b.double() ** 2)
    rel = (rr / bb).sqrt().cpu().numpy()
    its = info["iterations"].numpy()
    pair_rows = np.repeat(np.diff(off), 4).reshape(num_systems, 4)
    res = {"workload": f"{num_systems} 2D 5-point log-normal diffusion systems, sides uniform integers in "
                       f"[{lo},{hi}] (seed {seed}), sigma={sigma}, 4 rhs each, Jacobi-PCG to 1e-6 "
                       f"(FP32 matrix and Krylov vectors, compensated iterate returned in FP64)",
           "n_total": n, "nnz": nnz, "wall_s": float(min(times)), "device_s": info["seconds"],
           "converged_pairs": info["converged_pairs"], "total_pairs": info["total_pairs"],
           "max_iterations": info["max_iterations"], "median_iterations": float(np.median(its)),
           "graph_launches": info["graph_launches"], "tile_rows": info["tile_rows"],
           "rows_per_thread": info["rows_per_thread"], "verifications": info["verifications"],
           "diagonal_planes": info["diagonal_planes"], "symmetric_storage": info["symmetric_storage"],
           "max_true_relres_fp64": float(rel.max()),
           "row_iterations_per_s": float((pair_rows * its).sum() / min(times)),
           "setup": {"generate_s": t_gen, "coo_to_csr_s": t_csr, "first_solve_s_incl_tuning": t_first}}
    log("batched", res)
    return res


def fsai_section(dev, steps, warmup, reps, weights=None, only_headline=False):
    """Held-out 128^3 7-point log-normal diffusion (sigma 2): time to true relres
    1e-6 of plain FP32 CG and of FP64-refined CG / Jacobi-PCG / classical FSAI-PCG
    / learned FSAI-PCG (the stated 5-input network with FP32 and BF16 matrix
    values; the 8-feature variant with BF16 values), preconditioner set-up
    included; every error against a Jacobi-PCG FP64 oracle (torch only).
    Each solver: one warm-up, then the median of `reps` solves. The headline
    solve (stated-5 network, BF16 values) is then timed once more as `warmup`
    untimed + `steps` timed solves."""
    import numpy as np
    import torch

    from mpsparse_amd import learned_fsai as LF
    from mpsparse_amd import spmv

    side, sigma, seed = 128, 2.0, 2024
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    keys, vals, n = LF.diffusion_keys_torch((side, side, side), sigma, gen, dev)
    nnz = keys.numel()
    rp = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ci = torch.empty(nnz, dtype=torch.int32, device=dev)
    cv = torch.empty(nnz, dtype=torch.float32, device=dev)
    A = spmv.csr_tensor(keys, vals, rp, ci, cv, n, n)
    del keys, vals
    b64 = torch.randn(n, generator=gen, dtype=torch.float64, device=dev)
    b32 = b64.float()
    rows = torch.repeat_interleave(torch.arange(n, device=dev), (rp[1:] - rp[:-1]).long())
    cols = ci.long()
    bnorm = float(torch.linalg.vector_norm(b64))

    def relres64(x):
        ax = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(0, rows, cv.double() * x.double()[cols])
        return float(torch.linalg.vector_norm(b64 - ax)) / bnorm

    res = {"workload": f"diffusion 3D 7-point {side}^3 sigma={sigma} (held out: seed {seed}, side outside the "
                       f"training range)", "n": n, "nnz": nnz, "tol": 1e-6,
           "timing": f"per solver: 1 warm-up, then the median of {reps} solves (wall clock, set-up included)"}
    if only_headline:
        net = LF.load_network(weights or LF.DEFAULT_WEIGHTS, device=dev)
        x64 = torch.zeros(n, dtype=torch.float64, device=dev)

        def solve():
            F, _ = LF.build(A, net)
            x64.zero_()
            return LF.solve_refined(A, F, b64, x64, tol=1e-6, values="bf16")

        for _ in range(warmup):
            solve()
        torch.cuda.synchronize()
        ts = []
        for _ in range(steps):
            t1 = time.perf_counter()
            info = solve()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
        res["headline_steps"] = {"steps": steps, "median_s": float(np.median(ts)), "runs_s": ts,
                                 "inner": info["inner_iterations"], "outer": info["outer_iterations"],
                                 "true_relres_fp64": relres64(x64)}
        log("headline only", res["headline_steps"])
        return res
    t = time.perf_counter()
    op64 = stencil_operator64(rp, ci, cv, n, side)
    d64 = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(
        0, rows[rows == cols], cv[rows == cols].double())
    xo, ito = fp64_cg_torch(op64, b64, 1e-12, 100000, dinv=1.0 / d64)
    torch.cuda.synchronize()
    xo_norm = float(torch.linalg.vector_norm(xo))
    res["fp64_oracle"] = {"solver": "FP64 Jacobi-PCG, torch 7-plane stencil operator", "iterations": ito,
                          "wall_s": time.perf_counter() - t, "true_relres_fp64": relres64(xo)}
    log("fsai fp64 oracle", res["fp64_oracle"])

    def err(x):
        return float(torch.linalg.vector_norm(x.double() - xo)) / xo_norm

    def timed(fn):
        fn()  # warm-up
        times, out = [], None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), times, out

    # plain FP32 CG (the baseline)
    x32 = torch.zeros(n, device=dev)
    t_cg, runs, info = timed(lambda: A.cgd(b32, x32, tol=1e-6, maxiter=100000))
    res["plain_cg_fp32"] = {"wall_s": t_cg, "wall_s_runs": runs, "iterations": info["iterations"],
                            "status": info["status"], "true_relres_fp64": relres64(x32),
                            "rel_err_vs_fp64_oracle": err(x32)}
    log("plain cg", res["plain_cg_fp32"])
    x64 = torch.zeros(n, dtype=torch.float64, device=dev)

    def refined(precond):
        x64.zero_()
        return A.solve_refined(b64, x64, tol=1e-6, precond=precond)

    for name, pc in (("plain_cg_refined", "none"), ("jacobi_pcg_refined", "jacobi")):
        t_r, runs, rinfo = timed(lambda: refined(pc))
        res[name] = {"wall_s": t_r, "wall_s_runs": runs, "outer": rinfo["outer_iterations"],
                     "iterations": rinfo["inner_iterations"], "status": rinfo["status"],
                     "true_relres_fp64": relres64(x64), "rel_err_vs_fp64_oracle": err(x64)}
        log(name, res[name])

    nets = {"stated5": LF.load_network(weights or LF.DEFAULT_WEIGHTS, device=dev),
            "coupled8": LF.load_network(LF.COUPLED8_WEIGHTS, device=dev)}
    res["weights"] = {"stated5": os.path.relpath(weights or LF.DEFAULT_WEIGHTS, HERE),
                      "coupled8": os.path.relpath(LF.COUPLED8_WEIGHTS, HERE)}
    variants = (("classical_fsai_refined", "classical", "f32"),
                ("classical_fsai_refined_bf16", "classical", "bf16"),
                ("learned_fsai_stated5_refined_f32", "stated5", "f32"),
                ("learned_fsai_stated5_refined_bf16", "stated5", "bf16"),
                ("learned_fsai_coupled8_refined_bf16", "coupled8", "bf16"))

    def make_solver(kind, vt):
        holder = {}

        def setup():
            if kind == "classical":
                F = LF.fsai_tensor(A)
                holder["fallback_rows"] = F.classical()
            else:
                F, holder["nonpositive_diag"] = LF.build(A, nets[kind])
            holder["F"] = F
            return F

        def solve():
            F = setup()
            x64.zero_()
            return LF.solve_refined(A, F, b64, x64, tol=1e-6, values=vt)

        return holder, setup, solve

    for name, kind, vt in variants:
        holder, setup, solve = make_solver(kind, vt)
        t_s, _, _ = timed(setup)
        t_all, runs, rinfo = timed(solve)
        res[name] = {"wall_s": t_all, "wall_s_runs": runs, "setup_s": t_s, "outer": rinfo["outer_iterations"],
                     "iterations": rinfo["inner_iterations"], "fallbacks": rinfo["fallbacks"],
                     "status": rinfo["status"], "true_relres_fp64": relres64(x64),
                     "rel_err_vs_fp64_oracle": err(x64),
                     **{k: v for k, v in holder.items() if k != "F"}}
        if kind != "classical":
            res[name]["network"] = {"feature_set": kind, "inputs": "a_ij, a_ii, a_jj, row nnz, row |sum|"
                                    if kind == "stated5" else "the five above + s_low, s_all, 1/sqrt(1-s_low)"}
        log(name, res[name])
        holder.clear()

    # the headline solve, timed as the benchmark's steps
    _, setup, solve = make_solver("stated5", "bf16")
    for _ in range(warmup):
        solve()
    torch.cuda.synchronize()
    step_times = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        solve()
        torch.cuda.synchronize()
        step_times.append(time.perf_counter() - t1)
    res["headline_steps"] = {"steps": steps, "warmup": warmup, "total_s": time.perf_counter() - t0,
                             "median_s": float(np.median(step_times)), "runs_s": step_times}

    # the edge MLP on this matrix's lower pattern (stated-5 features)
    F, _ = LF.build(A, nets["stated5"])
    feat, _ = F.features()
    o = torch.empty(feat.shape[0], device=dev)
    ms = time_launches(lambda: LF._C.mlp_forward(nets["stated5"].params, feat, o), 50, 3)
    fpe = 2 * (8 * 64 + 64 * 64 + 64)  # the first layer runs K = 8 (5 features + 3 zero slots) on the MFMA
    res["mlp_forward"] = {"edges": int(feat.shape[0]), "ms": ms, "flops_per_edge": fpe,
                          "TFLOPs": feat.shape[0] * fpe / (ms * 1e-3) / 1e12}
    log("mlp forward", res["mlp_forward"])
    del F, feat, o
    # the dominant kernel of the headline step: q = A p on the 16-bit diagonal storage
    xa = torch.randn(n, device=dev)
    ya = torch.empty(n, device=dev)
    A.mv(xa, ya, "bf16")
    mv_ms = time_launches(lambda: A.mv(xa, ya, "bf16"), 100, 5)
    # the BF16 diagonals streamed, FP32 diagonal remainder, x, y
    mv_bytes = streamed_planes(A, 2, n) * 2 * n + 4 * n + 4 * n + 4 * n
    res["mv_bf16"] = {"ms": mv_ms, "algorithmic_bytes": mv_bytes, "GBps": mv_bytes / (mv_ms * 1e-3) / 1e9,
                      "note": "working set ~%d MB: resident in the 256 MB last-level cache, so not an HBM-bound "
                              "figure" % (mv_bytes // 2 ** 20)}
    log("mv bf16 128^3", res["mv_bf16"])
    base = min(res["plain_cg_fp32"]["wall_s"], res["plain_cg_refined"]["wall_s"])
    head = res["learned_fsai_stated5_refined_bf16"]["wall_s"]
    res["speedup_learned_bf16_vs_plain_cg"] = base / head
    res["speedup_learned_bf16_vs_jacobi_pcg"] = res["jacobi_pcg_refined"]["wall_s"] / head
    res["speedup_learned_bf16_vs_classical_fsai"] = res["classical_fsai_refined"]["wall_s"] / head
    res["speedup_learned_bf16_vs_classical_fsai_bf16"] = res["classical_fsai_refined_bf16"]["wall_s"] / head
    res["_csr"] = (rp, ci, cv, b32)  # for the CPU baseline; removed before printing
    return res


def direct_section(dev, sizes=(1000, 4000, 10000, 20000, 100000), nrhs=128):
    """The reference's direct-solver benchmark (sparse_bridge/benchmark.py:38-86):
    1D Laplacian + 0.1 I of each size, factorisation, then one solve with 128
    right-hand sides; scipy's sparse LU on the host next to it (the reference's
    own comparison) and ||X - X_splu|| (the reference passes below 1e-8). Plus
    a 2D 64x64 grid (band 64 after reordering) as a wider-band point."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from mpsparse_amd import problems
    from mpsparse_amd.sparse_qr import SparseQRSolver

    def case(name, cs, ri, v, n):
        A = sp.csc_matrix((v, ri, cs), shape=(n, n))
        tcs, tri, tv = (torch.from_numpy(a).to(dev) for a in (cs, ri, v))
        SparseQRSolver(tcs, tri, tv, n, n)  # warm-up (code paths, scratch)
        torch.cuda.synchronize()
        t = time.perf_counter()
        S = SparseQRSolver(tcs, tri, tv, n, n)
        torch.cuda.synchronize()
        t_fact = time.perf_counter() - t
        rng = np.random.default_rng(n)
        Bh = rng.random((n, nrhs))
        B = torch.from_numpy(Bh.T.copy()).to(dev).t()  # column-major, as the reference passes it
        S.solve(B)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            X = S.solve(B)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        t = time.perf_counter()
        lu = spla.splu(A)
        c_fact = time.perf_counter() - t
        t = time.perf_counter()
        Xc = lu.solve(Bh)
        c_solve = time.perf_counter() - t
        diff = float(np.linalg.norm(X.cpu().numpy() - Xc))
        r = {"n": n, "bandwidth": list(S.bandwidth), "nrhs": nrhs, "factor_ms": t_fact * 1e3,
             "solve_ms": float(np.median(ts)) * 1e3, "scipy_splu_factor_ms": c_fact * 1e3,
             "scipy_splu_solve_ms": c_solve * 1e3, "diff_vs_splu": diff, "pass": diff < 1e-8}
        log("direct", name, r)
        return r

    out = {"workload": "1D Laplacian + 0.1 I (sparse_bridge/benchmark.py generate_data), 128 right-hand sides; "
                       "FP64 banded Householder QR after reverse Cuthill-McKee", "cases": {}}
    for n in sizes:
        out["cases"][f"laplacian_1d_{n}"] = case(f"1d {n}", *problems.laplacian_1d_csc(n), n)
    side = 64
    T = sp.diags([-np.ones(side - 1), 4.1 * np.ones(side), -np.ones(side - 1)], [-1, 0, 1])
    A2 = (sp.kron(sp.eye(side), T) + sp.kron(sp.diags([-np.ones(side - 1), -np.ones(side - 1)], [-1, 1]),
                                              sp.eye(side))).tocsc()
    A2.sort_indices()
    out["cases"]["laplacian_2d_64x64"] = case("2d 64x64", A2.indptr.astype(np.int64), A2.indices.astype(np.int64),
                                              A2.data.astype(np.float64), side * side)
    return out


def precision_section(dev, empty, spmv_steps=50, only_products=False):
    """256^3 7-point log-normal diffusion: plain FP32 CG at 1e-6 next to FP64
    iterative refinement (FP32 / BF16 / FP16 values) to 1e-10."""
    import numpy as np
    import torch

    from mpsparse_amd import problems, spmv

    t0 = time.time()
    M = problems.diffusion_3d(256, sigma=1.0, seed=1)
    n = M.num_rows
    log(f"diffusion 256^3 n={n} nnz={M.nnz} generated in {time.time() - t0:.1f}s")
    rows_np = np.repeat(np.arange(n, dtype=np.uint64), np.diff(M.row_ptr.astype(np.int64)))
    keys = torch.from_numpy(problems.pack_keys(rows_np, M.col_ind).view(np.int64)).to(dev)
    del rows_np
    vals = torch.from_numpy(M.vals).to(dev)
    rp = torch.from_numpy(M.row_ptr.view(np.int32)).to(dev)
    ci = torch.from_numpy(M.col_ind.view(np.int32)).to(dev)
    A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
    del keys
    b64 = torch.from_numpy(problems.rhs(n, seed=1).astype(np.float64)).to(dev)
    b32 = b64.float()
    rows = torch.repeat_interleave(torch.arange(n, device=dev), (rp[1:] - rp[:-1]).long())
    cols = ci.long()

    def relres64(x):
        ax = torch.zeros(n, dtype=torch.float64, device=dev).index_add_(0, rows, vals.double() * x.double()[cols])
        return float(torch.linalg.vector_norm(b64 - ax) / torch.linalg.vector_norm(b64))

    res = {"workload": "diffusion_3d 256^3 sigma=1 seed=1 (7-point, SpMV alg bytes FP32 %d, BF16 %d)" % (
        M.nnz * 8 + 4 * (n + 1) + 8 * n, M.nnz * 6 + 4 * (n + 1) + 8 * n)}
    # the stencil product in diagonal storage: 7 value planes of n (zeros where
    # a diagonal leaves the grid), x, y (+ the FP32 diagonal remainder with 16-bit values)
    xm = torch.randn(n, device=dev)
    ym = torch.empty(n, device=dev)
    res["mv"] = {}
    for vt, vb, extra in (("f32", 4, 0), ("bf16", 2, 4 * n)):
        A.mv(xm, ym, vt)
        res["mv"]["storage"] = A.format  # known once the first product has probed the pattern
        ms = time_launches(lambda: A.mv(xm, ym, vt), spmv_steps, 5)
        alg = streamed_planes(A, vb, n) * vb * n + 4 * n + 4 * n + extra
        res["mv"][vt] = {"ms": ms, "algorithmic_bytes": alg, "GBps": alg / (ms * 1e-3) / 1e9,
                         "launches_timed": spmv_steps}
        log("stencil mv 256^3", vt, res["mv"][vt])
    del xm, ym
    # one CG step (FP32 values): a fixed 400-step run without convergence test
    xs = torch.zeros(n, device=dev)
    A.cgd(b32, xs, tol=0.0, maxiter=50, verify=False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    sinfo = A.cgd(b32, xs, tol=0.0, maxiter=400, verify=False)
    torch.cuda.synchronize()
    res["cg_step_us"] = {"f32": (time.perf_counter() - t) / max(1, sinfo["iterations"]) * 1e6,
                         "steps": sinfo["iterations"]}
    log("cg step 256^3", res["cg_step_us"])
    del xs
    if only_products:
        return res
    x32 = torch.zeros(n, device=dev)
    A.cgd(b32, x32, maxiter=20000)
    torch.cuda.synchronize()
    t = time.perf_counter()
    info = A.cgd(b32, x32, maxiter=20000)
    torch.cuda.synchronize()
    res["fp32_cg_1e-6"] = {"wall_s": time.perf_counter() - t, "iterations": info["iterations"],
                           "status": info["status"], "true_relres_fp64": relres64(x32)}
    log("fp32 cg", res["fp32_cg_1e-6"])
    t = time.perf_counter()
    op64 = stencil_operator64(rp, ci, vals, n, 256)
    xo, ito = fp64_cg_torch(op64, b64, 1e-11, 40000)
    torch.cuda.synchronize()
    res["fp64_oracle"] = {"solver": "FP64 CG, torch 7-plane stencil operator", "iterations": ito,
                          "wall_s": time.perf_counter() - t, "true_relres_fp64": relres64(xo)}
    log("fp64 oracle", res["fp64_oracle"])
    xnorm = float(torch.linalg.vector_norm(xo))
    res["fp32_cg_1e-6"]["rel_err_vs_fp64_oracle"] = float(torch.linalg.vector_norm(x32.double() - xo)) / xnorm
    for vt in ("f32", "bf16", "f16"):
        x64 = torch.zeros(n, dtype=torch.float64, device=dev)
        A.solve_refined(b64, x64, tol=1e-10, values=vt)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rinfo = A.solve_refined(b64, x64, tol=1e-10, values=vt)
        torch.cuda.synchronize()
        res[f"refined_{vt}_1e-10"] = {"wall_s": time.perf_counter() - t, "outer": rinfo["outer_iterations"],
                                      "inner_iterations": rinfo["inner_iterations"], "status": rinfo["status"],
                                      "fallbacks": rinfo["fallbacks"], "true_relres_fp64": relres64(x64),
                                      "rel_err_vs_fp64_oracle": float(torch.linalg.vector_norm(x64 - xo)) / xnorm}
        log("refined", vt, res[f"refined_{vt}_1e-10"])
    return res


if __name__ == "__main__":
    main()
