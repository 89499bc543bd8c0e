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
        "host": host_description(),
    }, y


def csr5_model_bytes(m, nnz):
    """(m + 1 + nnz) int32 indices and (2 nnz + m) doubles: values, x once per
    element, y (include/detail/utils.h:10-14)."""
    return (m + 1 + nnz) * 4 + (2 * nnz + m) * 8


def host_description():
    """The host the CPU leg ran on: CPU model, logical CPUs of the machine and of
    this process, NUMA nodes, and the OpenMP binding variables in effect."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and
                    d[4:].isdigit()])
    except OSError:
        numa = None
    return {"cpu_model": model, "logical_cpus": os.cpu_count(),
            "process_cpus": len(os.sched_getaffinity(0)), "numa_nodes": numa,
            "OMP_PLACES": os.environ.get("OMP_PLACES"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def gpu_clocks():
    """The active DPM level of the shader (sclk), memory (mclk) and fabric
    (fclk) clocks of each amdgpu device in sysfs (the starred line of
    pp_dpm_*), e.g. {"card1": {"sclk": "2400Mhz", ...}}; empty where the files
    are not readable."""
    out = {}
    base = "/sys/class/drm"
    try:
        cards = sorted(d for d in os.listdir(base) if d.startswith("card") and d[4:].isdigit())
    except OSError:
        return out
    for c in cards:
        e = {}
        for k in ("sclk", "mclk", "fclk"):
            try:
                with open(f"{base}/{c}/device/pp_dpm_{k}") as f:
                    for line in f:
                        if "*" in line:
                            e[k] = line.split(":", 1)[1].replace("*", "").strip()
            except OSError:
                pass
        if e:
            out[c] = e
    return out


def preheat(fn, floor_ms):
    """Runs fn back to back (a sync every 50 calls) for at least floor_ms of
    wall time, so the GPU leaves its idle clock state before the counted
    warm-up steps.  Returns what it ran."""
    import torch
    t0 = time.perf_counter()
    n = 0
    while floor_ms > 0:
        for _ in range(50):
            fn()
        n += 50
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= floor_ms:
            break
    return {"ms": round((time.perf_counter() - t0) * 1e3, 1), "launches": n}


def _time_calls(fn, reps=20, warmup=3, preheat_ms=200.0):
    """Median wall time (ms) of fn() on the current stream, by torch events
    around each call (fn launches on torch's current stream).  Untimed calls
    run first for at least preheat_ms, so the GPU is at its working clocks
    whatever ran before (host-only work lets it drop to idle clocks)."""
    import torch
    t0 = time.perf_counter()
    while True:
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= preheat_ms:
            break
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        ts.append((e0, e1))
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ts]))


def extras_ingest(A, rp, col, val, log_to):
    """Matrix ingestion on the device (sortCOORow + coo2csr's pointer step,
    src/spmv_helper.cu:277-573): the R-MAT matrix as COO in a random order,
    sorted back by (row, column) and compressed; checked against the
    generator's CSR.  Wall time of the synchronous calls (index validation and
    per-call scratch included).  CPU side: numpy's stable lexsort on a
    4M-element sample, one thread."""
    import torch
    from hpda_sblas_pnnl_amd.device import coo_ptr, coo_sort
    row = torch.repeat_interleave(torch.arange(A.m, dtype=torch.int32, device=rp.device),
                                  torch.diff(rp).long())
    g = torch.Generator(device=rp.device)
    g.manual_seed(5)
    perm = torch.randperm(A.nnz, device=rp.device, generator=g)
    times = []
    for _ in range(3):
        r, c, v = row[perm].contiguous(), col[perm].contiguous(), val[perm].contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        coo_sort(r, c, v, A.m, A.n)
        t1 = time.perf_counter()
        ptr = coo_ptr(r, A.m)
        t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t1))
    exact = bool(torch.equal(r, row) and torch.equal(c, col) and torch.equal(v, val) and
                 torch.equal(ptr, rp))
    ts, tp = min(times)
    k = min(A.nnz, 4_000_000)
    rr, cc = row[perm[:k]].cpu().numpy(), col[perm[:k]].cpu().numpy()
    t0 = time.perf_counter()
    np.lexsort((cc, rr))
    t_cpu = time.perf_counter() - t0
    out = {"nnz": A.nnz, "sort_ms": round(ts * 1e3, 2), "ptr_ms": round(tp * 1e3, 3),
           "Melem_per_s": round(A.nnz / (ts + tp) / 1e6, 1), "exact": exact,
           "cpu_sample": {"nnz": k, "ms": round(t_cpu * 1e3, 1),
                          "Melem_per_s": round(k / t_cpu / 1e6, 1), "cores": 1,
                          "kind": "numpy lexsort (stable)"}}
    log_to(f"[bench] ingest: {out}")
    return out


def extras_rmat(log_to):
    """R-MAT scale 22 (edge factor 17, binary values, alpha 1, beta 0.5): FP64
    kernels 1/2/3, COO and CSC forms; FP32 / BF16 values and a k=8 block of
    right-hand sides, each with its error against the FP64 oracle."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    from oracle import oracle as ora
    dev = torch.device("cuda:0")
    A = G.rmat_csr(scale=22, edge_factor=17, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    alpha, beta = 1.0, 0.5
    cores = len(os.sched_getaffinity(0))
    ref = ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0, threads=cores)
    rp, col, val = (torch.from_numpy(v).to(dev) for v in (A.row_ptr, A.col, A.val))
    dx, dy0 = torch.from_numpy(x).to(dev), torch.from_numpy(y0).to(dev)
    out = {"workload": "R-MAT scale 22, edge factor 17, dedup, binary values, beta 0.5",
           "m": A.m, "nnz": A.nnz, "cpu_baseline": cpu_sample(A, x, y0, alpha, beta)}
    log_to(f"[bench] rmat cpu baseline: {out['cpu_baseline']}")

    def run(plan, fmt_bytes_val=8, fmt="csr"):
        info = plan.info()
        dy = dy0.clone()
        plan.spmv(dx, dy, alpha, beta)
        torch.cuda.synchronize()
        err = ora.err_normwise(dy.cpu().numpy(), ref)
        maxrow = ora.err_maxrow(A.m, A.val, A.row_ptr, A.col, x, alpha, beta, y0,
                                dy.cpu().numpy(), ref)
        ms = _time_calls(lambda: plan.spmv(dx, dy, alpha, beta))
        cb = compulsory_bytes(A.m, A.n, A.nnz, fmt_bytes_val, 1, True)
        if fmt == "coo":
            cb += A.nnz * 4 - (A.m + 1) * 4
        return {"ms": round(ms, 4), "GFLOPs": round(2 * A.nnz / ms / 1e6, 1),
                "GBs": round(cb / ms / 1e6, 1), "frac": round(cb / ms / 1e6 / PEAK_HBM_GBS, 4),
                "err_normwise": err, "err_maxrow": maxrow, "kernel": info["kernel"],
                "variant": info["variant"]}

    kernels = {}
    for name, kw in (("auto", {}), ("kernel1_rowblock", {"kernel": "vector"}),
                     ("kernel1_segments", {"kernel": "vector", "variant": 6}),
                     ("kernel2_merge", {"kernel": "merge"}), ("kernel3_csr5", {"kernel": "csr5"})):
        p = SpmvPlan("csr", A.m, A.n, rp, col, val, **kw)
        kernels[name] = run(p)
        del p
        log_to(f"[bench] rmat {name}: {kernels[name]}")
    ri = torch.from_numpy(G.coo_rows(A)).to(dev)
    for name, kw in (("coo_auto", {}), ("coo_segmented", {"kernel": "coo"})):
        # AUTO on COO builds row pointers and runs the CSR chooser
        p = SpmvPlan("coo", A.m, A.n, ri, col, val, **kw)
        kernels[name] = run(p, fmt="coo")
        del p
    del ri
    cp, rows, cv = G.to_csc(A)
    dcp, drows, dcv = (torch.from_numpy(v).to(dev) for v in (cp, rows, cv))
    p = SpmvPlan("csc", A.m, A.n, dcp, drows, dcv)
    kernels["csc_transposed"] = run(p)
    del p, dcp, drows, dcv
    out["fp64"] = kernels
    base_ms = kernels["auto"]["ms"]  # narrow-value plans below are AUTO plans too
    out["ingest"] = extras_ingest(A, rp, col, val, log_to)

    prec = {}
    v32 = torch.from_numpy(A.val.astype(np.float32)).to(dev)
    p = SpmvPlan("csr", A.m, A.n, rp, col, v32)
    prec["fp32_values"] = run(p, 4)
    del p
    v16 = torch.from_numpy(G.to_bf16_bits(A.val).view(np.int16)).to(dev)
    p = SpmvPlan("csr", A.m, A.n, rp, col, v16)
    prec["bf16_values"] = run(p, 2)
    del p
    for k in prec:
        prec[k]["speedup_vs_fp64"] = round(base_ms / prec[k]["ms"], 3)
    # k = 8 right-hand sides, X row-major n x 8
    K = 8
    X = G.values(A.n * K, 0, 4, "uniform").reshape(A.n, K)
    Y0 = G.values(A.m * K, 0, 5, "uniform").reshape(A.m, K)
    dX, dY0 = torch.from_numpy(X).to(dev), torch.from_numpy(Y0).to(dev)
    p = SpmvPlan("csr", A.m, A.n, rp, col, val)
    dY = dY0.clone()
    p.spmm(dX, dY, alpha, beta)
    torch.cuda.synchronize()
    Yref = ora.csr_spmm(A.m, K, A.val, A.row_ptr, A.col, X, alpha, beta, Y0)
    err = ora.err_normwise(dY.cpu().numpy().ravel(), Yref.ravel())
    ms = _time_calls(lambda: p.spmm(dX, dY, alpha, beta))
    prec["spmm_k8_fp64"] = {"ms": round(ms, 4), "ms_per_vector": round(ms / K, 4),
                            "speedup_per_vector_vs_fp64": round(base_ms * K / ms, 3),
                            "GFLOPs": round(2 * A.nnz * K / ms / 1e6, 1), "err_normwise": err}
    del p, dX, dY, dY0
    out["precision"] = prec
    return out


def extras_batched(log_to, count=64):
    """64 matrices (log-uniform n in [1e4, 2e6], half banded, half R-MAT),
    FP32 values with FP64 accumulation, one persistent launch replayed as a
    hipGraph; against the same matrices run one FP64 plan at a time."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvBatch, SpmvPlan
    from oracle import oracle as ora
    dev = torch.device("cuda:0")
    t0 = time.time()
    mats = G.batched_workload(count, seed=0)
    total = sum(A.nnz for _, A in mats)
    log_to(f"[bench] batched: {count} matrices, nnz={total} generated in {time.time() - t0:.1f}s")
    cores = len(os.sched_getaffinity(0))
    items, plans, xs, refs, y64 = [], [], [], [], []
    for i, (kind, A) in enumerate(mats):
        x = G.values(A.n, i, 2, "uniform")
        dx = torch.from_numpy(x).to(dev)
        rp = torch.from_numpy(A.row_ptr).to(dev)
        col = torch.from_numpy(A.col).to(dev)
        v32 = torch.from_numpy(A.val.astype(np.float32)).to(dev)
        v64 = torch.from_numpy(A.val).to(dev)
        items.append(dict(m=A.m, n=A.n, row_ptr=rp, col=col, val=v32, x=dx,
                          y=torch.zeros(A.m, dtype=torch.float64, device=dev), alpha=1.0, beta=0.0))
        plans.append((SpmvPlan("csr", A.m, A.n, rp, col, v64), dx,
                      torch.zeros(A.m, dtype=torch.float64, device=dev), v64))
        refs.append(ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.0, np.zeros(A.m),
                                 threads=cores))
    batch = SpmvBatch(items)
    batch.graph_launch()
    torch.cuda.synchronize()
    errs = [ora.err_normwise(it["y"].cpu().numpy(), r) for it, r in zip(items, refs)]
    ms_graph = _time_calls(batch.graph_launch, reps=10)

    def fp64_all():
        for p, dx, dy, _ in plans:
            p.spmv(dx, dy)
    ms_fp64 = _time_calls(fp64_all, reps=10)
    errs64 = []
    for (p, dx, dy, _), r in zip(plans, refs):
        errs64.append(ora.err_normwise(dy.cpu().numpy(), r))
    choices = [batch.choice(i) for i in range(count)]
    cb32 = sum(compulsory_bytes(A.m, A.n, A.nnz, 4) for _, A in mats)
    out = {
        "workload": f"{count} matrices, n log-uniform in [1e4, 2e6] (seed 0), even banded "
                    "(half-width min(50000, n/40)), odd R-MAT edge factor 16 cut to n; "
                    "FP32 values, FP64 accumulation",
        "matrices": count, "nnz_total": total,
        "graph_ms": round(ms_graph, 3), "fp64_separate_ms": round(ms_fp64, 3),
        "speedup_vs_fp64_separate": round(ms_fp64 / ms_graph, 3),
        "GFLOPs": round(2 * total / ms_graph / 1e6, 1),
        "GBs": round(cb32 / ms_graph / 1e6, 1),
        "max_err_normwise": max(errs), "max_err_normwise_fp64_path": max(errs64),
        "shape": batch.shape(),
        "autotuner_subwaves": sorted({c["subwave"] for c in choices}),
    }
    return out


def extras_banded_precision(A, rp, col, dx, x, y0, alpha, beta, chunks, part, plan64, y_ref,
                            log_to):
    """The headline matrix with FP32 and BF16 values (FP64 x, y and
    accumulation), same chunks and partition: time per SpMV against the FP64
    plan and normwise / max-row error against the FP64 host oracle."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    from oracle import oracle as ora
    dev = torch.device("cuda:0")
    dy = torch.from_numpy(y0).to(dev)
    ms64 = _time_calls(lambda: plan64.spmv(dx, dy, alpha, beta), reps=20)
    out = {"fp64_ms": round(ms64, 4)}
    for name, vals, vb in (("fp32_values", A.val.astype(np.float32), 4),
                           ("bf16_values", G.to_bf16_bits(A.val).view(np.int16), 2)):
        dv = torch.from_numpy(vals).to(dev)
        p = SpmvPlan("csr", A.m, A.n, rp, col, dv, chunks=chunks, part_mode=part)
        dy = torch.from_numpy(y0).to(dev)
        p.spmv(dx, dy, alpha, beta)
        torch.cuda.synchronize()
        got = dy.cpu().numpy()
        ms = _time_calls(lambda: p.spmv(dx, dy, alpha, beta), reps=20)
        cb = compulsory_bytes(A.m, A.n, A.nnz, vb, 1, beta != 0)
        out[name] = {"ms": round(ms, 4), "GFLOPs": round(2 * A.nnz / ms / 1e6, 1),
                     "GBs": round(cb / ms / 1e6, 1), "frac": round(cb / ms / 1e6 / PEAK_HBM_GBS, 4),
                     "speedup_vs_fp64": round(ms64 / ms, 3), "variant": p.info()["variant"],
                     "err_normwise": ora.err_normwise(got, y_ref),
                     "err_maxrow": ora.err_maxrow(A.m, A.val, A.row_ptr, A.col, x, alpha, beta,
                                                  y0, got, y_ref)}
        log_to(f"[bench] banded {name}: {out[name]}")
        del p, dv
    return out


def extras_banded_formats(A, x, y0, alpha, beta, chunks, y_ref, log_to):
    """The headline matrix as row-sorted COO and as CSC (device arrays, plans
    with the same chunk count): kernel time per SpMV, GFLOP/s and roofline
    fraction on each format's own compulsory bytes (COO: a row index per
    element instead of row pointers; CSC: column pointers and row indices),
    error against the FP64 oracle."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    from oracle import oracle as ora
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    out = {}
    for fmt in ("coo", "csc"):
        if fmt == "coo":
            a0, a1, v = G.coo_rows(A), A.col, A.val
            cb = compulsory_bytes(A.m, A.n, A.nnz, 8, 1, beta != 0) + A.nnz * 4 - (A.m + 1) * 4
        else:
            a0, a1, v = G.to_csc(A)
            cb = compulsory_bytes(A.m, A.n, A.nnz, 8, 1, beta != 0) + (A.n - A.m) * 4
        d0, d1, dv = (torch.from_numpy(np.ascontiguousarray(t)).to(dev) for t in (a0, a1, v))
        p = SpmvPlan(fmt, A.m, A.n, d0, d1, dv, chunks=chunks)
        dy = torch.from_numpy(y0).to(dev)
        p.spmv(dx, dy, alpha, beta)
        torch.cuda.synchronize()
        err = ora.err_normwise(dy.cpu().numpy(), y_ref)
        ms = _time_calls(lambda: p.spmv(dx, dy, alpha, beta), reps=20)
        inf = p.info()
        out[fmt] = {"ms": round(ms, 4), "GFLOPs": round(2 * A.nnz / ms / 1e6, 1),
                    "GBs": round(cb / ms / 1e6, 1), "frac": round(cb / ms / 1e6 / PEAK_HBM_GBS, 4),
                    "err_normwise": err, "kernel": inf["kernel"], "variant": inf["variant"],
                    "chunks": chunks}
        log_to(f"[bench] banded {fmt}: {out[fmt]}")
        del p, d0, d1, dv, dy
        torch.cuda.empty_cache()
    return out


def cpu_sample(A, x, y0, alpha, beta, reps=10):
    """The oracle's row loop with OpenMP over rows on this host's cores:
    median of reps passes over the whole matrix (GFLOP/s), and one serial pass."""
    from oracle import oracle as ora
    cores = len(os.sched_getaffinity(0))
    ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0, threads=cores)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0, threads=cores)
        ts.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0)
    t1 = time.perf_counter() - t0
    t = float(np.median(ts))
    return {"value": round(2 * A.nnz / t / 1e9, 3), "unit": "GFLOP/s", "cores": cores,
            "kind": "port", "sample": f"whole matrix, median of {reps} OpenMP passes "
                                      f"({t * 1e3:.2f} ms each)",
            "serial_gflops": round(2 * A.nnz / t1 / 1e9, 3)}


def extras_powerlaw(log_to):
    """The n = 20,000 power-law matrix (alpha 1, beta 0), kernel chosen by the
    autotuner: time per SpMV (launch-bound: 4 MB of matrix stays in L2 / MALL)
    and the CPU baseline on the same matrix."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    from oracle import oracle as ora
    dev = torch.device("cuda:0")
    A = G.powerlaw_csr(20000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = np.zeros(A.m)
    ref = ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.0, y0)
    rp, col, val = (torch.from_numpy(v).to(dev) for v in (A.row_ptr, A.col, A.val))
    dx, dy = torch.from_numpy(x).to(dev), torch.from_numpy(y0).to(dev)
    p = SpmvPlan("csr", A.m, A.n, rp, col, val)
    p.spmv(dx, dy, 1.0, 0.0)
    torch.cuda.synchronize()
    err = ora.err_normwise(dy.cpu().numpy(), ref)
    ms = _time_calls(lambda: p.spmv(dx, dy, 1.0, 0.0), reps=200)
    inf = p.info()
    out = {"workload": "power-law n=20,000, p(L)~(L+2.23)^-2 on [1,2000], FP64 CSR",
           "nnz": A.nnz, "ms": round(ms, 4), "GFLOPs": round(2 * A.nnz / ms / 1e6, 1),
           "err_normwise": err, "kernel": inf["kernel"], "variant": inf["variant"],
           "cpu_baseline": cpu_sample(A, x, y0, 1.0, 0.0, reps=50)}
    # narrow values (FP64 x, y and sums) and k = 8 right-hand sides on the same matrix
    prec = {}
    for name, vals in (("fp32_values", A.val.astype(np.float32)),
                       ("bf16_values", G.to_bf16_bits(A.val).view(np.int16))):
        pv = SpmvPlan("csr", A.m, A.n, rp, col, torch.from_numpy(vals).to(dev))
        pv.spmv(dx, dy, 1.0, 0.0)
        torch.cuda.synchronize()
        e = ora.err_normwise(dy.cpu().numpy(), ref)
        msv = _time_calls(lambda: pv.spmv(dx, dy, 1.0, 0.0), reps=200)
        prec[name] = {"ms": round(msv, 4), "speedup_vs_fp64": round(ms / msv, 3),
                      "err_normwise": e, "variant": pv.info()["variant"]}
    k = 8
    X = G.values(A.n * k, 0, 4, "uniform").reshape(A.n, k)
    dX = torch.from_numpy(X).to(dev)
    dY = torch.zeros((A.m, k), dtype=torch.float64, device=dev)
    p.spmm(dX, dY, 1.0, 0.0)
    torch.cuda.synchronize()
    refk = ora.csr_spmm(A.m, k, A.val, A.row_ptr, A.col, X, 1.0, 0.0, np.zeros((A.m, k)))
    ek = ora.err_normwise(dY.cpu().numpy().ravel(), np.asarray(refk).ravel())
    msk = _time_calls(lambda: p.spmm(dX, dY, 1.0, 0.0), reps=200)
    prec["spmm_k8_fp64"] = {"ms": round(msk, 4), "ms_per_vector": round(msk / k, 4),
                            "speedup_per_vector_vs_fp64": round(ms * k / msk, 3),
                            "err_normwise": ek}
    out["precision"] = prec
    log_to(f"[bench] powerlaw: {out}")
    return out


def extras_dropin(A, x, y0, alpha, beta, log_to):
    """The reference's own entry point, spMV_mgpu_v1_numa (ngpu = 1, kernel 1),
    on host arrays: every call uploads the matrix, runs, and downloads y, so
    this is the PCIe-inclusive rate; the spmv_ret timings split it."""
    from hpda_sblas_pnnl_amd import spmv as S
    out = {}
    for rep in range(2):  # the first call also pays one-time device setup
        y = y0.copy()
        t0 = time.perf_counter()
        r = S.spMV_mgpu_v1_numa(A.m, A.n, A.nnz, alpha, A.val, A.row_ptr, A.col, x, beta, y,
                                ngpu=1, kernel=1)
        wall = time.perf_counter() - t0
        out = {"wall_ms": round(wall * 1e3, 2), "entry": "spMV_mgpu_v1_numa",
               "GFLOPs_pcie_inclusive": round(2 * A.nnz / wall / 1e9, 2),
               "part_ms": round(r.part_time * 1e3, 3), "comm_ms": round(r.comm_time * 1e3, 3),
               "comp_ms": round(r.comp_time * 1e3, 3), "merg_ms": round(r.merg_time * 1e3, 3)}
    log_to(f"[bench] drop-in spMV_mgpu_v1_numa: {out}")
    out.update(extras_v2_stream(A, x, y0, alpha, beta, log_to))
    return out


def extras_v2_stream(A, x, y0, alpha, beta, log_to):
    """spMV_mgpu_v2 with the matrix streamed from host memory in nine tasks
    over four streams (copies of one task overlap kernels of the others),
    beside the in-core v2 drop-in at the same tasks; PCIe-inclusive."""
    from hpda_sblas_pnnl_amd import spmv as S
    out = {}
    nb = (A.nnz + 8) // 9
    for mode, key in (("1", "v2_streamed"), ("0", "v2_in_core")):
        os.environ["SBLAS_V2_STREAM"] = mode
        try:
            for rep in range(2):
                y = y0.copy()
                t0 = time.perf_counter()
                r = S.spMV_mgpu_v2(A.m, A.n, A.nnz, alpha, A.val, A.row_ptr, A.col, x, beta, y,
                                   ngpu=1, kernel=1, nb=nb, copy_of_workspace=4)
                wall = time.perf_counter() - t0
        finally:
            del os.environ["SBLAS_V2_STREAM"]
        out[key] = {"wall_ms": round(wall * 1e3, 2), "tasks": 9, "streams": 4,
                    "GFLOPs_pcie_inclusive": round(2 * A.nnz / wall / 1e9, 2),
                    "comm_ms": round(r.comm_time * 1e3, 3), "comp_ms": round(r.comp_time * 1e3, 3),
                    "host_to_device_GBs": round((A.nnz * 12 + (A.m + 1) * 4) / wall / 1e9, 1)}
        log_to(f"[bench] drop-in spMV_mgpu_v2 {key}: {out[key]}")
    return out


# Sources of the column-panel kernel: the measured HBM traffic recorded in
# profiles/traffic_<workload>.json holds their hash, and a changed kernel makes
# the recorded figure stale.
KERNEL_SOURCES = ["hpda_sblas_pnnl_amd/csrc/sblas_panel.hip",
                  "hpda_sblas_pnnl_amd/csrc/panel_device.hpp",
                  "hpda_sblas_pnnl_amd/csrc/internal.hpp"]


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed counter
    runs (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel's readings), or None
    when no record exists or the kernel's sources changed since it was taken."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        rec = json.load(f)
    if rec.get("source_hash") != kernel_source_hash():
        log(f"[bench] {p}: recorded for other kernel sources; traffic not reported")
        return None
    return rec.get("traffic_bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="banded", choices=["banded", "rmat", "powerlaw"])
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--chunks", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0,
                    help="kernel 1 form: 1 sub-wave, 2 row blocks, 3 column panels, 4 sorted "
                         "blocks, 6 row blocks over column segments")
    ap.add_argument("--values", default="f64", choices=["f64", "f32", "bf16"],
                    help="value type of the timed plan (profiling the narrow forms)")
    ap.add_argument("--items", type=int, default=0)
    ap.add_argument("--subwave", type=int, default=0)
    ap.add_argument("--xcd-remap", type=int, default=-1)
    ap.add_argument("--preheat-ms", type=float, default=500.0,
                    help="untimed launches for at least this long before the warm-up steps")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the R-MAT precision and batched measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device copy first (reference for HBM byte counters)")
    ap.add_argument("--only", choices=["ingest", "stream", "batch"],
                    help="run one side measurement alone and print its JSON (for profiling): "
                         "device COO ingestion of R-MAT scale 22, the streamed v2 drop-in "
                         "on the banded matrix, or the 64-matrix batch")
    args = ap.parse_args()

    if args.gpus != 1:
        log(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
            "design is described but not built")
        sys.exit(2)

    import torch
    from hpda_sblas_pnnl_amd.device import SpmvPlan

    assert torch.cuda.is_available(), "no HIP device"
    dev = torch.device("cuda:0")
    if args.only == "ingest":
        from hpda_sblas_pnnl_amd import gen as G
        A = G.rmat_csr(scale=22, edge_factor=17, seed=0)
        rp, col, val = (torch.from_numpy(v).to(dev) for v in (A.row_ptr, A.col, A.val))
        print(json.dumps({"ingest": extras_ingest(A, rp, col, val, log)}), flush=True)
        return
    if args.only == "stream":
        A, x, y0, alpha, beta, _, _, _ = make_workload("banded")
        print(json.dumps({"stream": extras_v2_stream(A, x, y0, alpha, beta, log)}), flush=True)
        return
    if args.only == "batch":
        print(json.dumps({"batched": extras_batched(log)}), flush=True)
        return
    if args.copy_probe:
        # a plain copy of known size: reads 1 GiB, writes 1 GiB
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy probe: 1073741824 bytes read, 1073741824 bytes written")
        del a, b
    A, x, y0, alpha, beta, chunks, part, desc = make_workload(args.workload)
    if args.chunks:
        chunks = args.chunks
    rp = torch.from_numpy(A.row_ptr).to(dev)
    col = torch.from_numpy(A.col).to(dev)
    if args.values == "f64":
        val = torch.from_numpy(A.val).to(dev)
    elif args.values == "f32":
        val = torch.from_numpy(A.val.astype(np.float32)).to(dev)
    else:
        from hpda_sblas_pnnl_amd import gen as G
        val = torch.from_numpy(G.to_bf16_bits(A.val).view(np.int16)).to(dev)
    val_bytes = {"f64": 8, "f32": 4, "bf16": 2}[args.values]
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

    clocks_idle = gpu_clocks()
    heat = preheat(lambda: plan.spmv(dx, dy, alpha, beta), args.preheat_ms)
    for _ in range(args.warmup):
        plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    clocks_before = gpu_clocks()
    plan.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    clocks_after = gpu_clocks()
    ktimes = plan.kernel_times_ms()
    plan.set_timing(False)
    log(f"[bench] preheat {heat}; {args.steps} steps in {wall * 1e3:.2f} ms; clocks "
        f"idle {clocks_idle} before {clocks_before} after {clocks_after}")

    ms_step = wall * 1e3 / args.steps
    gflops = 2.0 * A.nnz / (ms_step * 1e-3) / 1e9
    cb = compulsory_bytes(A.m, A.n, A.nnz, val_bytes, 1, beta != 0)
    k_ms = float(np.mean(ktimes)) if len(ktimes) else float("nan")
    achieved = cb / (k_ms * 1e-3) / 1e9
    kstats = ({"min": round(float(np.min(ktimes)), 4), "median": round(float(np.median(ktimes)), 4),
               "max": round(float(np.max(ktimes)), 4), "launches": int(len(ktimes))}
              if len(ktimes) else {})
    log(f"[bench] kernel ms per launch: mean {k_ms:.4f} {kstats}")
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
        "dtype": "f64" if args.values == "f64" else f"{args.values} values, f64 accumulate",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "m": A.m, "n": A.n, "nnz": A.nnz, "alpha": alpha,
                   "beta": beta, "chunks": info["chunks"], "kernel": info["kernel"],
                   "subwave": info["subwave"], "variant": info["variant"],
                   "items": info["items"], "xcd_remap": info["xcd_remap"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(args.workload),
                     "algorithmic_bytes": cb, "kernel_ms": round(k_ms, 4),
                     "kernel_ms_stats": kstats},
        # untimed: launches run before the counted warm-up steps so the GPU
        # leaves its idle clocks; clocks as sysfs reports them around the loop
        "preheat": heat,
        "gpu_clocks": {"before": clocks_before, "after": clocks_after},
        # the CSR5 library's own byte model (include/detail/utils.h:10-14), which
        # reads x once per element: an upper-bound view, not the roofline
        "csr5_model": {"bytes": csr5_model_bytes(A.m, A.nnz),
                       "GBs": round(csr5_model_bytes(A.m, A.nnz) / (k_ms * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu:
        cpu, y_ref = cpu_baseline(A, x, y0, alpha, beta)
        from oracle import oracle as ora
        line["cpu_baseline"] = cpu
        line["err_normwise_vs_oracle"] = ora.err_normwise(y_gpu, y_ref)
    if not args.no_extras and args.workload == "banded":
        extras = {}
        from oracle import oracle as ora
        y_ref = ora.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, alpha, beta, y0,
                             threads=len(os.sched_getaffinity(0)))
        try:
            extras["banded_precision"] = extras_banded_precision(
                A, rp, col, dx, x, y0, alpha, beta, chunks, part, plan, y_ref, log)
        except Exception as e:  # noqa: BLE001  (the main line must still print)
            extras["banded_precision"] = {"error": f"{type(e).__name__}: {e}"}
        del plan, rp, col, val
        torch.cuda.empty_cache()
        try:
            extras["banded_formats"] = extras_banded_formats(A, x, y0, alpha, beta, chunks, y_ref,
                                                             log)
        except Exception as e:  # noqa: BLE001
            extras["banded_formats"] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
        try:
            extras["powerlaw"] = extras_powerlaw(log)
        except Exception as e:  # noqa: BLE001
            extras["powerlaw"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            extras["dropin_host_arrays"] = extras_dropin(A, x, y0, alpha, beta, log)
        except Exception as e:  # noqa: BLE001
            extras["dropin_host_arrays"] = {"error": f"{type(e).__name__}: {e}"}
        for name, fn in (("rmat", extras_rmat), ("batched", extras_batched)):
            t0 = time.time()
            try:
                extras[name] = fn(log)
            except Exception as e:  # noqa: BLE001  (the main line must still print)
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] extras {name} done in {time.time() - t0:.1f}s")
            torch.cuda.empty_cache()
        line["extras"] = extras
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
