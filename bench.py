#!/usr/bin/env python3
"""Benchmark of the Tensors sparse hot path on one MI355X.

Workload (one "step"): the random-CSR SpMM suite -- n = 1,000,000 rows/cols, row
lengths Poisson(32) clipped to [1, 512], distinct uniform columns per row, N(0,1)
values, dense RHS uniform[-1,1]; Y <- A X (alpha = 1, beta = 0) for nrhs 1, 4, 16, 64
in FP64 and in FP32 (8 SpMM launches).  value = algorithmic bytes of the step / step
time (GB/s), with A, X and Y resident in HBM before the timed region.

Algorithmic bytes per SpMM: nnz*(s_v + 4) + 4*(m+1) + n*k*s + m*k*s (beta = 0).

Prints one JSON line.  Options: --gpus N (only 1 is supported), --steps K, --warmup W,
--no-cpu (skip the CPU baseline leg).
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

METRIC = "SpMM GB/s & % HBM roofline; CG solves/s + p99 latency at 1e-10 residual"
PEAK_HBM_GBPS = 8000.0
M = N = 1_000_000
KS = (1, 4, 16, 64)
SEED = 2002


def alg_bytes(nnz, m, n, k, sv):
    return nnz * (sv + 4) + 4 * (m + 1) + n * k * sv + m * k * sv


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cg", action="store_true", help="skip the CG (configs 3) section")
    ap.add_argument("--cg-grid", type=int, default=128)
    ap.add_argument("--cg-reps", type=int, default=2)
    return ap.parse_args()


def read_profile_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprof PMC summary
    (profiles/spmm_pmc_random_csr_1e6.txt, if present): (fetch+write bytes, note)."""
    path = os.path.join(ROOT, "profiles", "spmm_pmc_random_csr_1e6.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch"):
                k, v = line.split("=", 1)
                vals[k.strip()] = float(v.split()[0])
    return vals.get("traffic_bytes_per_launch")


def cpu_baseline(rp, ci, v64, v32, X64, X32, budget_s):
    """The oracle (same source, reference release flags) on the host cores, on the same
    matrix; each case repeated until the time budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle  # test-side checker; here only as the timed CPU baseline

    threads = min(16, os.cpu_count() or 1)
    nnz = int(rp[-1])
    total_bytes, total_time, reps_done = 0.0, 0.0, 0
    per_case_budget = budget_s / (2 * len(KS))
    for dtype, vals, Xall in ((np.float64, v64, X64), (np.float32, v32, X32)):
        sv = 8 if dtype == np.float64 else 4
        for k in KS:
            X = np.ascontiguousarray(Xall[: N * k].reshape(N, k))
            Y = np.zeros((M, k), dtype=dtype)
            oracle.spmm(rp, ci, vals, M, N, 1.0, X, k, 0.0, Y, k, k, NRHS=k, threads=threads, fast=True)
            t0 = time.perf_counter()
            reps = 0
            while True:
                oracle.spmm(rp, ci, vals, M, N, 1.0, X, k, 0.0, Y, k, k, NRHS=k, threads=threads, fast=True)
                reps += 1
                if time.perf_counter() - t0 > per_case_budget:
                    break
            dt = (time.perf_counter() - t0) / reps
            total_bytes += alg_bytes(nnz, M, N, k, sv)
            total_time += dt
            reps_done += reps
    return {
        "value": round(total_bytes / total_time / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the same 8 SpMMs (FP64/FP32 x nrhs 1,4,16,64) on the full n=1e6 matrix, "
                  f"{reps_done} timed calls in ~{budget_s:.0f}s, fixed-NRHS path, -O3 -ffast-math",
    }


def cg_section(args, dev):
    """Single SPD solve: 3D 7-point Laplacian on a g^3 grid + 1e-2 shift, 32 RHS
    uniform[-1,1] (seed 3003), to 1e-10 relative residual: plain FP64 CG (the
    ConjugateGradient recurrences, P = I, |r.r| <= tol^2 |b.b|) vs FP32-inner CG with
    FP64 refinement (true ||b - A x|| / ||b|| <= 1e-10).  Reports solves/s, per-solve
    latencies, iterations, true residuals and the mixed result's error vs FP64."""
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd import generators as gen
    from tensors_amd.solvers import ConjugateGradient, cg_solve_mixed
    from tensors_amd.sparse import MatrixCSR, spmm_raw

    g, k, tol = args.cg_grid, 32, 1e-10
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    A = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    B = torch.from_numpy(gen.uniform(3003, n * k).reshape(n, k)).to(dev)
    rpd, cid, vd = A.device_arrays(torch.float64)
    nnz = A.NonzeroCount()

    def true_res(X):
        R = B.clone()
        spmm_raw(n, n, nnz, rpd, cid, vd, -1.0, X, k, 1.0, R, k, k)
        return (R.norm(dim=0) / B.norm(dim=0)).max().item()

    cg = ConjugateGradient(n, 5000, k)
    X64 = torch.zeros(n, k, dtype=torch.float64, device=dev)
    cg(A, None, 1.0, B, 0.0, X64, tol)  # warm-up (also allocates)
    lat64 = []
    for _ in range(args.cg_reps):
        X64.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok64 = cg(A, None, 1.0, B, 0.0, X64, tol)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t0)
    it64 = cg.IterationCount()
    res64 = true_res(X64)
    Xm = torch.zeros(n, k, dtype=torch.float64, device=dev)
    _, a32 = None, A.device_arrays(torch.float32)[2]
    cg_solve_mixed(A, B, Xm, tol=tol, a32=a32)  # warm-up
    latm = []
    for _ in range(args.cg_reps):
        Xm.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        okm, outer, inner, rr = cg_solve_mixed(A, B, Xm, tol=tol, a32=a32)
        torch.cuda.synchronize()
        latm.append(time.perf_counter() - t0)
    resm = true_res(Xm)
    err = ((Xm - X64).abs().max() / X64.abs().max()).item()
    # algorithmic bytes of one FP64 CG iteration (P = I): SpMM + 11 vector passes
    spmm_b = nnz * 12 + 4 * (n + 1) + 2 * n * k * 8
    it_bytes = spmm_b + 11 * n * k * 8
    t64 = float(np.median(lat64))
    tm = float(np.median(latm))
    return {
        "workload": f"3D 7-point Laplacian {g}^3 + 1e-2 shift, n={n}, nnz={nnz}, 32 RHS uniform[-1,1], "
                    f"tol {tol} relative residual, P=I",
        "fp64": {"solves_per_s": round(1.0 / t64, 3), "ms_per_solve": round(t64 * 1e3, 2),
                 "latencies_ms": [round(x * 1e3, 2) for x in lat64], "iterations": it64, "succeeded": bool(ok64),
                 "true_rel_residual": res64,
                 "GBps_per_iteration": round(it_bytes * it64 / t64 / 1e9, 1)},
        "mixed": {"solves_per_s": round(1.0 / tm, 3), "ms_per_solve": round(tm * 1e3, 2),
                  "latencies_ms": [round(x * 1e3, 2) for x in latm], "outer": outer, "inner_iterations": inner,
                  "succeeded": bool(okm), "true_rel_residual": resm, "error_vs_fp64": err},
        "speedup_mixed_over_fp64": round(t64 / tm, 3),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd import generators as gen
    from tensors_amd.sparse import spmm_raw

    dev = torch.device("cuda:0")
    rp, ci, v64, v32 = gen.random_csr(M, N, 32.0, 1, 512, SEED, want64=True, want32=True)
    nnz = int(rp[-1])
    kmax = max(KS)
    X64h = gen.uniform(SEED, N * kmax)
    X32h = X64h.astype(np.float32)
    rp32 = torch.from_numpy(rp.astype(np.int32)).to(dev)
    cid = torch.from_numpy(ci).to(dev)
    vals = {np.float64: torch.from_numpy(v64).to(dev), np.float32: torch.from_numpy(v32).to(dev)}
    Xs = {np.float64: torch.from_numpy(X64h).to(dev), np.float32: torch.from_numpy(X32h).to(dev)}
    cases = []
    for dtype in (np.float64, np.float32):
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        sv = 8 if dtype == np.float64 else 4
        for k in KS:
            X = Xs[dtype][: N * k].view(N, k)
            Y = torch.empty(M, k, dtype=tdt, device=dev)
            cases.append({"dtype": "f64" if sv == 8 else "f32", "k": k, "X": X, "Y": Y, "A": vals[dtype],
                          "bytes": alg_bytes(nnz, M, N, k, sv)})
    stream = torch.cuda.current_stream()

    def launch(c):
        return spmm_raw(M, N, nnz, rp32, cid, c["A"], 1.0, c["X"], c["k"], 0.0, c["Y"], c["k"], c["k"])

    for _ in range(args.warmup):
        for c in cases:
            launch(c)
    torch.cuda.synchronize()
    # timed region: events around every launch on the launch stream
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cases]
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        for i, c in enumerate(cases):
            ev[s][i][0].record(stream)
            launch(c)
            ev[s][i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_case_ms = [float(np.mean([ev[s][i][0].elapsed_time(ev[s][i][1]) for s in range(args.steps)]))
                   for i in range(len(cases))]
    step_bytes = sum(c["bytes"] for c in cases)
    ms_per_step = wall * 1e3 / args.steps
    kern_ms = sum(per_case_ms)
    value = step_bytes / (kern_ms * 1e-3) / 1e9

    # measured streaming bandwidth (copy and read) on a 2 GiB buffer
    nb = 2 << 30
    src = torch.empty(nb // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    sink = torch.zeros(1, dtype=torch.float32, device=dev)
    for _ in range(2):
        capi.call("tg_probe_copy", src.data_ptr(), dst.data_ptr(), nb, stream.cuda_stream)
        capi.call("tg_probe_read", src.data_ptr(), nb, sink.data_ptr(), stream.cuda_stream)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record(stream)
    for _ in range(5):
        capi.call("tg_probe_copy", src.data_ptr(), dst.data_ptr(), nb, stream.cuda_stream)
    e[1].record(stream)
    for _ in range(5):
        capi.call("tg_probe_read", src.data_ptr(), nb, sink.data_ptr(), stream.cuda_stream)
    e[2].record(stream)
    torch.cuda.synchronize()
    copy_gbps = 2 * nb * 5 / (e[0].elapsed_time(e[1]) * 1e-3) / 1e9
    read_gbps = nb * 5 / (e[1].elapsed_time(e[2]) * 1e-3) / 1e9
    del src, dst

    dom = int(np.argmax(per_case_ms))
    dc = cases[dom]
    achieved = dc["bytes"] / (per_case_ms[dom] * 1e-3) / 1e9
    traffic = read_profile_traffic()
    breakdown = [{"dtype": c["dtype"], "nrhs": c["k"], "ms": round(t, 4),
                  "GBps": round(c["bytes"] / (t * 1e-3) / 1e9, 1),
                  "frac_peak": round(c["bytes"] / (t * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4),
                  "frac_measured_read": round(c["bytes"] / (t * 1e-3) / 1e9 / read_gbps, 4)}
                 for c, t in zip(cases, per_case_ms)]
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "random CSR SpMM suite: n=1e6, Poisson(32) rows clipped [1,512], "
                               "FP64 and FP32, nrhs 1/4/16/64, alpha=1 beta=0, int32 indices",
                   "nnz": nnz, "nrhs": list(KS)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic,
                     "kernel": f"spmm {dc['dtype']} nrhs={dc['k']}",
                     "alg_bytes_per_launch": dc["bytes"]},
        "hbm_measured": {"copy_GBps": round(copy_gbps, 1), "read_GBps": round(read_gbps, 1)},
        "breakdown": breakdown,
    }
    # free the SpMM suite's buffers before the solver section
    del cases, vals, Xs, rp32, cid
    torch.cuda.empty_cache()
    if not args.no_cg:
        out["cg"] = cg_section(args, dev)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(rp, ci, v64, v32, X64h, X32h, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
