#!/usr/bin/env python3
"""SpMV benchmark on one MI355X.

Default workload (one JSON line on stdout): the banded FEM-like FP64 CSR
matrix, n = 2,000,000, row lengths round(N(140, 40)) clipped to [8, 400]
(~2.8e8 nonzeros), columns uniform within +-50,000 of the diagonal, values
N(0, 1), x U(-1, 1), alpha = 1, beta = 0, nnz-partitioned into 8 chunks with
the two-level rule of src/dspmv_mgpu_v1_numa.cu (one node) and split rows
merged on the device.  A step is one y = alpha*A*x + beta*y over the whole
matrix with A, x, y resident in HBM.

  value        GFLOP/s = 2*nnz / (wall time per step)
  roofline     compulsory bytes of the SpMV kernel / its mean duration
               (HIP events around each launch), against 8 TB/s
  cpu_baseline the oracle's row loop (test/dspmv_test_smt.cu:39-52) with
               OpenMP over rows on this host's cores, same matrix

Other workloads: --workload rmat (R-MAT scale 22, edge factor 17, beta 0.5)
or powerlaw (n = 20,000).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def compulsory_bytes(m, n, nnz, val_bytes=8, k=1, beta_nonzero=False):
    """Bytes an SpMV must move: values + column indices once, row pointers,
    x once, y written (and read when beta != 0)."""
    b = nnz * (val_bytes + 4) + (m + 1) * 4 + n * k * 8 + m * k * 8
    if beta_nonzero:
        b += m * k * 8
    return b


def make_workload(name):
    from hpda_sblas_pnnl_amd import gen as G
    t0 = time.time()
    if name == "banded":
        A = G.banded_csr(2_000_000, seed=0)
        alpha, beta, chunks, part = 1.0, 0.0, 8, "v1_numa"
        desc = ("banded FEM-like n=2,000,000, rows N(140,40) clipped [8,400], "
                "cols within +-50,000, FP64 CSR, 8 chunks (v1_numa rule), device merge")
    elif name == "rmat":
        A = G.rmat_csr(scale=22, edge_factor=17, seed=0)
        alpha, beta, chunks, part = 1.0, 0.5, 1, "v1"
        desc = "R-MAT scale 22 (a,b,c,d=.57,.19,.19,.05), edge factor 17, dedup, binary, FP64 CSR"
    elif name == "powerlaw":
        A = G.powerlaw_csr(20000, seed=0)
        alpha, beta, chunks, part = 1.0, 0.0, 1, "v1"
        desc = "power-law n=20,000, p(L)~(L+2.23)^-2 on [1,2000], FP64 CSR"
    else:
        raise SystemExit(f"unknown workload {name}")
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform") if beta != 0 else np.zeros(A.m)
    log(f"[bench] generated {name}: m={A.m} nnz={A.nnz} in {time.time() - t0:.1f}s")
    return A, x, y0, alpha, beta, chunks, part, desc


def cpu_baseline(A, x, y0, alpha, beta, budget_s=20.0):
    """The oracle's host loop with OpenMP over rows (bitwise equal to the
    serial loop), timed on this host; also returns its y for the error check."""
    from oracle import oracle as ora
    cores = len(os.sched_getaffinity(0))
    # warm-up (first touch), then timed repetitions within the budget
    y = ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0, threads=cores)
    reps, t_used = 0, 0.0
    times = []
    while t_used < budget_s * 0.5 and reps < 10:
        t0 = time.perf_counter()
        ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0, threads=cores)
        dt = time.perf_counter() - t0
        times.append(dt)
        t_used += dt
        reps += 1
    t_all = float(np.median(times))
    # one serial pass (the reference's own host path is serial)
    t0 = time.perf_counter()
    ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0)
    t_serial = time.perf_counter() - t0
    gflops = 2.0 * A.nnz / t_all / 1e9
    return {
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"whole matrix, median of {reps} OpenMP passes over rows "
                   f"({cores} threads, {t_all * 1e3:.1f} ms each); one serial pass "
                   f"{t_serial * 1e3:.1f} ms = {2.0 * A.nnz / t_serial / 1e9:.3f} GFLOP/s"),
        "serial_gflops": round(2.0 * A.nnz / t_serial / 1e9, 3),
    }, y


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("traffic_bytes")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="banded", choices=["banded", "rmat", "powerlaw"])
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--chunks", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="kernel 1 form: 1 sub-wave, 2 row blocks")
    ap.add_argument("--items", type=int, default=0)
    ap.add_argument("--subwave", type=int, default=0)
    ap.add_argument("--xcd-remap", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()

    if args.gpus != 1:
        log(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
            "design is described but not built")
        sys.exit(2)

    import torch
    from hpda_sblas_pnnl_amd.device import SpmvPlan

    assert torch.cuda.is_available(), "no HIP device"
    dev = torch.device("cuda:0")
    A, x, y0, alpha, beta, chunks, part, desc = make_workload(args.workload)
    if args.chunks:
        chunks = args.chunks
    rp = torch.from_numpy(A.row_ptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    val = torch.from_numpy(A.val).to(dev)
    dx = torch.from_numpy(x).to(dev)
    dy0 = torch.from_numpy(y0).to(dev)
    dy = dy0.clone()
    kernel = int(args.kernel) if args.kernel.isdigit() else args.kernel
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel=kernel, chunks=chunks, part_mode=part,
                    variant=args.variant, items=args.items, subwave=args.subwave,
                    xcd_remap=args.xcd_remap)
    info = plan.info()
    log(f"[bench] plan: {info}")

    # correctness of one step against the oracle (error reported in the line)
    plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    y_gpu = dy.cpu().numpy()

    for _ in range(args.warmup):
        plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    plan.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ktimes = plan.kernel_times_ms()
    plan.set_timing(False)
    log(f"[bench] {args.steps} steps in {wall * 1e3:.2f} ms")

    ms_step = wall * 1e3 / args.steps
    gflops = 2.0 * A.nnz / (ms_step * 1e-3) / 1e9
    cb = compulsory_bytes(A.m, A.n, A.nnz, 8, 1, beta != 0)
    k_ms = float(np.mean(ktimes)) if len(ktimes) else float("nan")
    achieved = cb / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "FP64 SpMV GFLOP/s + % HBM roofline",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "m": A.m, "n": A.n, "nnz": A.nnz, "alpha": alpha,
                   "beta": beta, "chunks": info["chunks"], "kernel": info["kernel"],
                   "subwave": info["subwave"], "variant": info["variant"],
                   "items": info["items"], "xcd_remap": info["xcd_remap"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(args.workload),
                     "algorithmic_bytes": cb, "kernel_ms": round(k_ms, 4)},
    }
    if not args.no_cpu:
        cpu, y_ref = cpu_baseline(A, x, y0, alpha, beta)
        from oracle import oracle as ora
        line["cpu_baseline"] = cpu
        line["err_normwise_vs_oracle"] = ora.err_normwise(y_gpu, y_ref)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
