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


_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def alg_bytes(nnz, m, n, k, sv):
    return nnz * (sv + 4) + 4 * (m + 1) + n * k * sv + m * k * sv


def line_fill_bytes(nnz, m, k, sv, line=128):
    """Traffic when no X line is reused from cache: every nonzero fills the whole 128-byte
    lines of its k-value X row, plus A and Y once.  For the random-column matrix with X far
    above the L2s and the Infinity Cache (nrhs 16-64) this is the floor the gathers set."""
    return nnz * (-(-k * sv // line)) * line + nnz * (sv + 4) + 4 * (m + 1) + m * k * sv


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=20, help="timed calls per CPU baseline case (median)")
    ap.add_argument("--cpu-warmup", type=int, default=3, help="untimed calls per CPU baseline case")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG (configs 3) section")
    ap.add_argument("--cg-grid", type=int, default=128)
    ap.add_argument("--cg-reps", type=int, default=2)
    ap.add_argument("--cpu-cg-iters", type=int, default=20, help="oracle CG iterations timed on the host")
    ap.add_argument("--cg-fused", type=int, default=-1,
                    help="CG iteration form (tg_tune_set cg_fused): 0 four kernels, 1 p.u in the SpMM, "
                         "2 fused direction, -1 library default")
    ap.add_argument("--cg-xdefer", type=int, default=-1,
                    help="deferral depth of the CG x update (tg_tune_set cg_xdefer): 1, 2 or 4, -1 library default")
    ap.add_argument("--gmres-delayed", type=int, default=-1,
                    help="GMRES step (tg_tune_set gmres_delayed): 1 second Gram-Schmidt pass delayed to the next "
                         "step (two basis passes), 0 plain CGS2 (three)")
    ap.add_argument("--cg-panel", type=int, default=-1,
                    help="rows of the CG SpMM (tg_tune_set cg_panel): 1 panel order where the band allows, "
                         "0 contiguous chunks, -1 library default")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven CG section")
    ap.add_argument("--batch-systems", type=int, default=2000)
    ap.add_argument("--batch-iters", type=int, default=16, help="CG iterations enqueued per poll")
    ap.add_argument("--batch-tile-rows", type=int, default=0, help="rows per workgroup tile (0 = library default)")
    ap.add_argument("--batch-stream", type=int, default=-1, help="batched SpMM kernel: 1 streaming, 0 row per lane")
    ap.add_argument("--batch-replace-drop-pct", type=int, default=-1,
                    help="residual replacement when the estimate fell to this percent (default 10)")
    ap.add_argument("--batch-check-every", type=int, default=-1,
                    help="iterations between residual-replacement checkpoints (default 16)")
    ap.add_argument("--batch-graph", type=int, default=-1,
                    help="batched CG steps as captured hipGraphs (1 default, 0 eager launches)")
    ap.add_argument("--batch-spmm", type=int, default=-1,
                    help="SpMM of the batched engine: 0 row loop, 1 pipelined row loop, 2 LDS-staged indices (default)")
    ap.add_argument("--batch-xcd-order", type=int, default=-1,
                    help="batched engine tile list: 0 round-robin over the XCDs (default), 1 contiguous runs per XCD")
    ap.add_argument("--batch-idx16", type=int, default=-1,
                    help="16-bit column offsets in the batched SpMM (1 default, 0 int32)")
    ap.add_argument("--batch-split", type=int, default=-1,
                    help="batched iteration: 1 direction pass + SpMM of p, 0 direction rebuilt in the SpMM")
    ap.add_argument("--batch-seq-sample", type=int, default=32,
                    help="systems solved one at a time by the drop-in FP64 CG for comparison")
    ap.add_argument("--no-split", action="store_true",
                    help="do not register the values with the plan (no re-laid slice layout)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuners (library default SpMM shapes and batch tiles)")
    ap.add_argument("--slices", type=int, default=-1,
                    help="column slices of the planned SpMM (tg_tune_set spmm_slices; -1 = library default)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop section")
    ap.add_argument("--line-rate", action="store_true",
                    help="add the line-rate probe: the FP64 16-column SpMM of the suite's row lengths with the "
                         "columns limited to [0, W) for X of 3.8-122 MB")
    ap.add_argument("--long-rows", action="store_true",
                    help="add the long-row probe: planned FP64 SpMM on 12 M nonzeros spread evenly or with 2 rows of "
                         "1e6 nonzeros, the long-row split on and off")
    ap.add_argument("--no-bsr", action="store_true", help="skip the block-CSR SpMM section")
    ap.add_argument("--bsr-vector", type=int, default=-1,
                    help="block-CSR lane mapping (tg_tune_set bsr_vector: 1 vector, 0 scalar; -1 default)")
    ap.add_argument("--no-mm", action="store_true", help="skip the MatrixMarket I/O section")
    ap.add_argument("--no-gmres", action="store_true", help="skip the single-system GMRES section")
    ap.add_argument("--serve-gmres-workers", type=int, default=6,
                    help="GMRES worker threads (own streams) of the serving loop")
    ap.add_argument("--serve-gmres-batch-rows", type=int, default=131072,
                    help="GMRES requests of up to this many rows join the batched engine (larger: worker streams)")
    ap.add_argument("--serve-gmres-batch", type=int, default=1,
                    help="1: GMRES requests join the batched GMRES engine; 0: one worker stream per request")
    ap.add_argument("--serve-gmres-mixed", type=int, default=1,
                    help="1: GMRES requests on the worker streams run the mixed solver (FP32 cycles, FP64 "
                         "refinement; tg_tune_set serve_gmres_mixed, the default); 0: tg_gmres_solve")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--serve-rates", type=str, default="0.25,0.5,1,2",
                    help="arrival rates as multiples of the measured saturated service rate")
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


def host_cpus():
    """Threads of the CPU baseline and a description of the host: the process's CPU share
    (its affinity mask, capped by OMP_NUM_THREADS where the box sets it) and what lscpu
    says about sockets / cores / threads."""
    try:
        mask = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        mask = os.cpu_count() or 1
    share = mask
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    desc = {"nproc": os.cpu_count(), "affinity_cpus": mask, "threads_used": share}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            key = key.strip()
            if key in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                desc[key] = val.strip()
    except Exception:
        pass
    return share, desc


def cpu_baseline(rp, ci, v64, v32, X64, X32, reps, warmup):
    """The oracle (same source, the reference's release flags) on the host's CPU share,
    threads pinned, on the same matrix: per case `warmup` untimed calls, then the median
    of `reps` timed calls (steady clock)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle  # test-side checker; here only as the timed CPU baseline

    threads, desc = host_cpus()
    oracle.lib(True).orc_set_pinning(1)
    nnz = int(rp[-1])
    total_bytes, total_time = 0.0, 0.0
    cases = []
    for dtype, vals, Xall in ((np.float64, v64, X64), (np.float32, v32, X32)):
        sv = 8 if dtype == np.float64 else 4
        for k in KS:
            X = np.ascontiguousarray(Xall[: N * k].reshape(N, k))
            Y = np.zeros((M, k), dtype=dtype)
            for _ in range(warmup):
                oracle.spmm(rp, ci, vals, M, N, 1.0, X, k, 0.0, Y, k, k, NRHS=k, threads=threads, fast=True)
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                oracle.spmm(rp, ci, vals, M, N, 1.0, X, k, 0.0, Y, k, k, NRHS=k, threads=threads, fast=True)
                times.append(time.perf_counter() - t0)
            dt = float(np.median(times))
            b = alg_bytes(nnz, M, N, k, sv)
            total_bytes += b
            total_time += dt
            cases.append({"dtype": "f64" if sv == 8 else "f32", "nrhs": k, "median_ms": round(dt * 1e3, 2),
                          "GBps": round(b / dt / 1e9, 2)})
            log(f"CPU baseline {cases[-1]}")
    oracle.lib(True).orc_set_pinning(0)
    return {
        "value": round(total_bytes / total_time / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the same 8 SpMMs (FP64/FP32 x nrhs 1,4,16,64) on the full n=1e6 matrix; per case {warmup} "
                  f"warm-up calls, then the median of {reps} timed calls; {threads} pinned threads (the "
                  f"process's CPU share); fixed-NRHS path, -O3 -ffast-math",
        "host": desc,
        "cases": cases,
    }


def cpu_cg_sample(g, k, iters):
    """The oracle CG (reference release flags, the host's CPU share, threads pinned, the
    vector updates threaded as ConjugateGradient.hpp:262-335 does) on the CG section's
    system: a 3-iteration warm-up call, then one timed call of `iters` iterations;
    returns (seconds per iteration, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle  # test-side checker; here only as the timed CPU baseline

    from tensors_amd import generators as gen

    threads, _ = host_cpus()
    oracle.lib(True).orc_set_pinning(1)
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    B = gen.uniform(3003, n * k).reshape(n, k)
    X = np.zeros((n, k))
    oracle.cg(n, rp, ci, val, B, X, k, 3, 1e-30, threads=threads, fast=True)  # warm-up
    t0 = time.perf_counter()
    oracle.cg(n, rp, ci, val, B, X, k, iters, 1e-30, threads=threads, fast=True)
    dt = (time.perf_counter() - t0) / iters
    oracle.lib(True).orc_set_pinning(0)
    return dt, threads


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

    if args.cg_fused >= 0:
        capi.call("tg_tune_set", b"cg_fused", args.cg_fused)
    if args.cg_panel >= 0:
        capi.call("tg_tune_set", b"cg_panel", args.cg_panel)
    if args.cg_xdefer >= 0:
        capi.call("tg_tune_set", b"cg_xdefer", args.cg_xdefer)
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
        log(f"CG fp64 solve {lat64[-1]:.3f}s")
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
        log(f"CG mixed solve {latm[-1]:.3f}s")
    resm = true_res(Xm)
    err = ((Xm - X64).abs().max() / X64.abs().max()).item()
    # bytes of one FP64 CG iteration as executed (P = I): SpMM (p read once, u written) +
    # 8 vector passes (r update: r, u read, r written; p step: r, p, x read, p, x written)
    spmm_b = nnz * 12 + 4 * (n + 1) + 2 * n * k * 8
    it_bytes = spmm_b + 8 * n * k * 8
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
        "cpu": cpu_cg_part(args, g, k, it64),
    }


def cpu_cg_part(args, g, k, it64):
    if args.no_cpu:
        return None
    per_it, threads = cpu_cg_sample(g, k, args.cpu_cg_iters)
    return {"seconds_per_iteration": round(per_it, 4), "cores": threads, "kind": "port",
            "est_solve_s": round(per_it * it64, 2), "est_solves_per_s": round(1.0 / (per_it * it64), 4),
            "sample": f"oracle CG (FP64, P = I, -O3 -ffast-math, SpMM and vector updates threaded over "
                      f"{threads} pinned threads) on the same system: a 3-iteration warm-up call, then "
                      f"{args.cpu_cg_iters} iterations timed in one call; solve time extrapolated to the GPU's "
                      f"{it64} iterations"}


def batch_specs(count, seed=4004):
    """Uneven SPD systems: n log-uniform in [1e3, 2e5], 2D or 3D weighted grid Laplacian
    (log-normal(0,1) edge coefficients, diagonal shift 1e-3), 1-8 RHS each."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n_t = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        dims = 2 if rng.random() < 0.5 else 3
        side = max(2, int(round(n_t ** (1.0 / dims))))
        specs.append((dims, side, int(rng.integers(1, 9))))
    return specs


def batch_section(args, dev):
    """Batched uneven solves to 1e-10: every system added at t0 to the batched CG engine
    (Jacobi), FP64 recurrences vs FP32-inner CG with FP64 refinement.  Reports solves/s,
    p50/p99 completion latency, iterations and the true residuals reached; the drop-in
    FP64 CG solving a sample of the systems one at a time is timed beside it."""
    from concurrent.futures import ThreadPoolExecutor

    import torch

    from tensors_amd import generators as gen
    from tensors_amd.batch import BatchCG
    from tensors_amd.solvers import ConjugateGradient
    from tensors_amd.sparse import MatrixCSR, spmm_raw

    specs = batch_specs(args.batch_systems)
    tol = 1e-10
    from tensors_amd import _capi as capi

    if args.batch_tile_rows:
        capi.call("tg_tune_set", b"batch_tile_rows", args.batch_tile_rows)
    if args.batch_stream >= 0:
        capi.call("tg_tune_set", b"batch_stream", args.batch_stream)
    if args.batch_split >= 0:
        capi.call("tg_tune_set", b"batch_split", args.batch_split)
    if args.batch_idx16 >= 0:
        capi.call("tg_tune_set", b"batch_idx16", args.batch_idx16)
    if args.batch_graph >= 0:
        capi.call("tg_tune_set", b"batch_graph", args.batch_graph)
    if args.batch_xcd_order >= 0:
        capi.call("tg_tune_set", b"batch_xcd_order", args.batch_xcd_order)
    if args.batch_spmm >= 0:
        capi.call("tg_tune_set", b"batch_spmm", args.batch_spmm)
    if args.batch_replace_drop_pct > 0:
        capi.call("tg_tune_set", b"batch_replace_drop_pct", args.batch_replace_drop_pct)
    if args.batch_check_every > 0:
        capi.call("tg_tune_set", b"batch_check_every", args.batch_check_every)

    def make(i):
        d, side, k = specs[i]
        return gen.weighted_grid(d, (side,) * d, 1e-3, seed=9000 + i)

    t_gen = time.perf_counter()
    with ThreadPoolExecutor(8) as ex:
        host = list(ex.map(make, range(len(specs))))
    t_gen = time.perf_counter() - t_gen
    log(f"batch: generated {len(host)} systems in {t_gen:.1f}s")
    g = torch.Generator(device=dev)
    g.manual_seed(4004)
    systems = []
    for (rp, ci, val), (_, _, k) in zip(host, specs):
        n = len(rp) - 1
        rpd = torch.from_numpy(rp.astype(np.int32)).to(dev)
        cid = torch.from_numpy(ci).to(dev)
        vd = torch.from_numpy(val).to(dev)
        rows = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int32), rpd[1:] - rpd[:-1])
        dg = torch.zeros(n, dtype=torch.float64, device=dev)
        m = cid == rows
        dg[rows[m].long()] = vd[m]
        B = torch.rand(n, k, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        systems.append({"n": n, "k": k, "nnz": len(ci), "rp": rpd, "ci": cid, "val": vd, "d": 1.0 / dg, "B": B})
    del host
    torch.cuda.synchronize()
    total_rows = sum(s["n"] for s in systems)
    total_nnz = sum(s["nnz"] for s in systems)

    def true_res(s, X):
        R = s["B"].clone()
        spmm_raw(s["n"], s["n"], s["nnz"], s["rp"], s["ci"], s["val"], -1.0, X, s["k"], 1.0, R, s["k"], s["k"])
        return (R.norm(dim=0) / s["B"].norm(dim=0)).max().item()

    def run(mode):
        """mode 0: FP64, the reference's stop test; 1: mixed; 2: FP64 with residual replacement and
        the true-residual stop (the same accuracy as mixed)."""
        mixed = mode == 1
        Xs = [torch.zeros(s["n"], s["k"], dtype=torch.float64, device=dev) for s in systems]
        torch.cuda.synchronize()
        bt = BatchCG()
        t0 = time.perf_counter()
        ids = {}
        for i, s in enumerate(systems):
            ids[bt.add(s["rp"], s["ci"], s["val"], s["B"], Xs[i], tol, 20000, diag=s["d"], mixed=mixed,
                       true_residual=mode == 2)] = i
        t_add = time.perf_counter() - t0
        done_t = {}
        for sid in bt.finished():
            done_t[sid] = time.perf_counter() - t0
        last = 0.0
        while True:
            active = bt.active()[0]
            if active == 0:
                break
            bt.step(args.batch_iters)
            now = time.perf_counter() - t0
            for sid in bt.finished():
                done_t[sid] = now
            if now - last > 10.0:
                log(f"batch mode {mode}: {active} systems active at {now:.1f}s")
                last = now
        wall = time.perf_counter() - t0
        res = {sid: bt.result(sid) for sid in ids}
        bt.close()
        lat = np.array(sorted(done_t.values()))
        tr = np.array([true_res(systems[i], Xs[i]) for sid, i in ids.items()])
        iters = np.array([res[sid]["iterations"] for sid in ids])
        # bytes of the executed schedule per iteration: nnz*(s_v + 4) + 4(n+1) + 2*n*s_v (Jacobi
        # diagonal, read by the direction and the update pass) + 10*n*k*s (direction: r, p_old,
        # x read, p, x written; SpMM: p read once, u written; update: u, r read, r written);
        # a residual replacement adds an FP64 SpMV over x (x64 + x32 for mixed) plus b read and
        # r written (and the x fold for mixed)
        sv = 4 if mixed else 8
        alg = 0.0
        for sid, i in ids.items():
            s_ = systems[i]
            n_, k_, z_ = s_["n"], s_["k"], s_["nnz"]
            alg += res[sid]["iterations"] * (z_ * (sv + 4) + 4 * (n_ + 1) + 2 * n_ * sv + 10 * n_ * k_ * sv)
            if mixed:
                alg += res[sid]["outer"] * (z_ * 12 + 4 * (n_ + 1) + n_ * k_ * (8 + 12 + 8 + 4 + 12))
            elif mode == 2:
                alg += res[sid]["outer"] * (z_ * 12 + 4 * (n_ + 1) + n_ * k_ * (8 + 8 + 8))
        return Xs, {
            "solves_per_s": round(len(systems) / wall, 2), "wall_s": round(wall, 3), "submit_s": round(t_add, 3),
            "alg_GB": round(alg / 1e9, 1), "alg_GBps": round(alg / wall / 1e9, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "succeeded": int(sum(r["succeeded"] for r in res.values())),
            "iterations_mean": round(float(iters.mean()), 1), "iterations_max": int(iters.max()),
            "outer_mean": round(float(np.mean([r["outer"] for r in res.values()])), 2),
            "true_rel_residual_max": float(tr.max()),
            "frac_true_rel_residual_le_1e-10": round(float((tr <= 1e-10).mean()), 4)}

    log(f"batch: {len(systems)} systems resident, {total_rows} rows, {total_nnz} nnz")
    tuned = None
    if not args.no_tune and not args.batch_tile_rows and args.batch_spmm < 0:
        # the project's autotuner picks the engine's tile height and SpMM form on a sample
        # of the workload (every 20th system, mixed mode, a fixed number of iterations)
        from tensors_amd import autotune

        sample = [(s["rp"], s["ci"], s["val"], s["d"], s["B"]) for s in systems[::20]]
        t_tune = time.perf_counter()
        best, table = autotune.tune_batch(sample, iters=16, warm=2, mixed=1)
        tuned = {"chosen": best, "sample_systems": len(sample), "tune_s": round(time.perf_counter() - t_tune, 2),
                 "ms_per_iteration": [dict(c, ms=round(t, 3)) for c, t in table]}
        log(f"batch: tuned {best} in {tuned['tune_s']}s")
    X64, r64 = run(0)
    log(f"batch fp64: {r64}")
    Xt, rt = run(2)
    log(f"batch fp64, true-residual stop: {rt}")
    Xm, rm = run(1)
    log(f"batch mixed: {rm}")
    rm["max_error_vs_fp64"] = max(((a - b).abs().max() / b.abs().max()).item() for a, b in zip(Xm, X64))
    rm["max_error_vs_fp64_true_residual"] = max(((a - b).abs().max() / b.abs().max()).item() for a, b in zip(Xm, Xt))
    del Xm, X64, Xt
    # the drop-in FP64 CG, one system at a time, on a sample
    sample = systems[: max(0, args.batch_seq_sample)]
    seq = None
    if sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j, s in enumerate(sample):
            if j % 8 == 0:
                log(f"batch: sequential drop-in CG system {j}/{len(sample)}")
            A = MatrixCSR.from_device(s["rp"], s["ci"], s["val"], s["n"], s["n"])
            X = torch.zeros(s["n"], s["k"], dtype=torch.float64, device=dev)
            ConjugateGradient(s["n"], 20000, s["k"])(A, s["d"], 1.0, s["B"], 0.0, X, tol)
        torch.cuda.synchronize()
        tseq = time.perf_counter() - t0
        seq = {"systems": len(sample), "solves_per_s": round(len(sample) / tseq, 2),
               "note": "the same first systems of the batch, tg_cg_solve one at a time"}
    return {
        "workload": f"{len(systems)} SPD systems, n log-uniform [1e3, 2e5] (total rows {total_rows}, nnz {total_nnz}), "
                    "2D/3D grid Laplacians with log-normal(0,1) edge coefficients + 1e-3 shift, 1-8 RHS "
                    "uniform[-1,1], Jacobi, tol 1e-10; all added at t0",
        "generation_s": round(t_gen, 2),
        "fp64_batched": r64, "fp64_true_residual_batched": rt, "mixed_batched": rm,
        "speedup_mixed_over_fp64_batched": round(rm["solves_per_s"] / r64["solves_per_s"], 3),
        "speedup_mixed_over_fp64_at_equal_accuracy": round(rm["solves_per_s"] / rt["solves_per_s"], 3),
        "speedup_mixed_batched_over_fp64_dropin_sequential": (round(rm["solves_per_s"] / seq["solves_per_s"], 3)
                                                                if seq else None),
        "note": "fp64_batched stops on the reference's preconditioned test (|r.z| <= tol^2 |b.Pb|); "
                "fp64_true_residual_batched and mixed_batched stop when every column has the true "
                "||b - A x|| / ||b|| <= 1e-10",
        "fp64_dropin_sequential_sample": seq,
        "autotuned_engine_shape": tuned,
    }


def serve_requests(count, seed=5005):
    """Request mix: half SPD CG requests shaped like the batched systems (mixed, Jacobi,
    1e-10), half 2D upwind convection-diffusion GMRES(30) requests (n log-uniform in
    [1e4, 1e6], Peclet log-uniform in [1, 100], P = I, 1e-8, 100 restarts)."""
    import scipy.sparse  # noqa: F401  (only to keep import cost out of the timed loop)

    from tensors_amd import generators as gen

    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        if rng.random() < 0.5:
            n_t = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
            dims = 2 if rng.random() < 0.5 else 3
            side = max(2, int(round(n_t ** (1.0 / dims))))
            k = int(rng.integers(1, 9))
            rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=seed * 10 + s)
            n = len(rp) - 1
            rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
            dg = np.zeros(n)
            m = ci == rows
            dg[rows[m]] = val[m]
            B = gen.uniform(seed * 100 + s, n * k).reshape(n, k)
            out.append(("cg", rp.astype(np.int32), ci, val, B, 1.0 / dg))
        else:
            n_t = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
            g = max(2, int(round(np.sqrt(n_t))))
            pe = float(np.exp(rng.uniform(0.0, np.log(100.0))))
            rp, ci, val = gen.convection_diffusion_2d(g, pe)
            n = g * g
            B = gen.uniform(seed * 100 + s, n).reshape(n, 1)
            out.append(("gmres", rp.astype(np.int32), ci, val, B, None))
    return out


def serve_section(args, dev):
    """Serving loop: Poisson arrivals of a CG/GMRES request mix into tg_server (200 GB
    admission budget, host arrays in and out, so latency includes the PCIe transfers).
    First every request at once (saturated service rate), then Poisson arrivals at the
    requested multiples of that rate; p50/p99 latency = completion - arrival."""
    from tensors_amd import _capi as capi
    from tensors_amd.serving import Server

    capi.call("tg_tune_set", b"serve_gmres_batch", int(args.serve_gmres_batch))
    capi.call("tg_tune_set", b"serve_gmres_batch_rows", int(args.serve_gmres_batch_rows))
    capi.call("tg_tune_set", b"serve_gmres_mixed", int(args.serve_gmres_mixed))
    t_gen = time.perf_counter()
    reqs = serve_requests(args.serve_requests)
    t_gen = time.perf_counter() - t_gen
    log(f"serve: generated {len(reqs)} requests in {t_gen:.1f}s")
    budget = 200 << 30

    def submit(srv, r):
        kind, rp, ci, val, B, d = r
        if kind == "cg":
            return srv.submit_cg(rp, ci, val, B, tol=1e-10, max_iter=20000, inv_diag=d, mixed=True)
        return srv.submit_gmres(rp, ci, val, B, tol=1e-8, restart=30, max_restarts=100)

    def run(rate, seed):
        srv = Server(budget, args.serve_gmres_workers, args.batch_iters)
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate > 0 else np.zeros(len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]
        ids, results = [], {}
        t0 = time.perf_counter()
        last = t0
        for r, a in zip(reqs, arrivals):
            while True:
                now = time.perf_counter() - t0
                if now >= a:
                    break
                for sid in srv.wait(min(0.002, a - now)):
                    results[sid] = srv.result(sid)
            ids.append(submit(srv, r))
            if time.perf_counter() - last > 10:
                log(f"serve rate {rate:.2f}/s: {len(ids)} submitted, {len(results)} done")
                last = time.perf_counter()
        while len(results) < len(ids):
            for sid in srv.wait(0.05):
                results[sid] = srv.result(sid)
            if time.perf_counter() - last > 10:
                log(f"serve rate {rate:.2f}/s: draining, {len(results)}/{len(ids)} done")
                last = time.perf_counter()
        st = srv.stats()
        srv.close()
        res = [results[i] for i in ids]
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        t_first = min(r["t_submit"] for r in res)
        t_last = max(r["t_done"] for r in res)
        kinds = [q[0] for q in reqs]
        ok = [r["succeeded"] and r["status"] == 0 for r in res]
        lat_k = {k: lat[[i for i, q in enumerate(kinds) if q == k]] for k in ("cg", "gmres")}
        return {
            "offered_rate_per_s": None if rate <= 0 else round(rate, 3),
            "solves_per_s": round(len(res) / (t_last - t_first), 3),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "cg_p50_ms": round(float(np.percentile(lat_k["cg"], 50)) * 1e3, 1) if len(lat_k["cg"]) else None,
            "gmres_p50_ms": round(float(np.percentile(lat_k["gmres"], 50)) * 1e3, 1) if len(lat_k["gmres"]) else None,
            "succeeded": int(sum(ok)), "requests": len(res),
            "peak_admitted_GB": round(st["peak_bytes"] / 1e9, 3),
        }

    sat = run(0.0, 5005)
    log(f"serve saturated: {sat}")
    sweep = []
    for mult in [float(x) for x in args.serve_rates.split(",") if x]:
        rate = mult * sat["solves_per_s"]
        sweep.append(dict(run(rate, 5005 + len(sweep) + 1), load=mult))
        log(f"serve load {mult}: {sweep[-1]}")
    n_cg = sum(1 for q in reqs if q[0] == "cg")
    return {
        "workload": f"{len(reqs)} requests: {n_cg} SPD CG (weighted 2D/3D grids, n log-uniform [1e3, 2e5], "
                    f"1-8 RHS, Jacobi, FP32 inner + FP64 refinement to 1e-10) and {len(reqs) - n_cg} GMRES(30) "
                    "(2D upwind convection-diffusion, n log-uniform [1e4, 1e6], Peclet log-uniform [1, 100], "
                    f"P = I, 1e-8, 100 restarts); host arrays in/out; 200 GB admission budget; GMRES "
                    + (f"requests of up to {args.serve_gmres_batch_rows} rows in the batched GMRES engine, larger "
                       f"ones on {args.serve_gmres_workers} worker streams" if args.serve_gmres_batch else
                       f"requests on {args.serve_gmres_workers} worker streams")
                    + (" (mixed GMRES: FP32 cycles, FP64 refinement)" if args.serve_gmres_mixed else ""),
        "generation_s": round(t_gen, 2),
        "saturated": sat,
        "poisson_sweep": sweep,
    }


def line_rate_section(dev):
    """What bounds the random-column SpMM: the suite's rows (Poisson(32), 1e6 rows) with the
    columns drawn from [0, W), FP64, 16 right-hand sides (one 128-byte line of X per
    nonzero), plain tg_csr_spmm.  With X small enough for the L2s every gather hits there,
    so the time left is the rate at which the CUs issue random line requests; at the suite's
    width X sits in the Infinity Cache.  Reports per W: time, line requests per second
    (one per nonzero plus A and Y) and the rate in bytes."""
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd import generators as gen

    s = torch.cuda.current_stream()
    k, out = 16, []
    for W in (31250, 125000, 500000, 1000000):
        rp, ci, v64, _ = gen.random_csr(M, W, 32.0, 1, 512, SEED, want64=True, want32=False)
        nnz = int(rp[-1])
        rpd = torch.from_numpy(rp.astype(np.int32)).to(dev)
        cid = torch.from_numpy(ci).to(dev)
        vd = torch.from_numpy(v64).to(dev)
        X = torch.from_numpy(gen.uniform(SEED + 7, W * k)).to(dev).view(W, k)
        Y = torch.empty(M, k, dtype=torch.float64, device=dev)

        def go():
            capi.call("tg_csr_spmm", M, W, nnz, rpd.data_ptr(), capi.I32, cid.data_ptr(), vd.data_ptr(), capi.F64,
                      1.0, X.data_ptr(), k, 0.0, Y.data_ptr(), k, k, k, None, 0, s.cuda_stream)

        for _ in range(3):
            go()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(10):
            go()
        e1.record(s)
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) * 1e-3 / 10
        fill = line_fill_bytes(nnz, M, k, 8)
        out.append({"W": W, "X_MB": round(W * k * 8 / 2 ** 20, 1), "nnz": nnz, "us": round(sec * 1e6, 1),
                    "G_lines_per_s": round(fill / 128 / sec / 1e9, 1), "line_GBps": round(fill / sec / 1e9, 1)})
        log(f"line rate: {out[-1]}")
        del rpd, cid, vd, X, Y
        torch.cuda.empty_cache()
    return {"workload": "1e6 rows, Poisson(32) row lengths clipped [1,512], columns uniform in [0, W), FP64, "
                        "16 right-hand sides (one 128 B line of X per nonzero), tg_csr_spmm", "cases": out}


def long_rows_section(dev):
    """Planned FP64 SpMM (MatrixCSR.Dot's path) on 1e6 x 1e6 matrices of about 12 M
    nonzeros: row lengths Poisson(12) clipped to [1, 64], or Poisson(10) plus 2 rows of
    1e6 nonzeros; columns uniform; nrhs 1, 4, 16; the plan's long-row split (rows of more
    than 8192 nonzeros) on and off (knob spmm_long_row)."""
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd.sparse import SlicePlan

    s = torch.cuda.current_stream()
    rng = np.random.default_rng(SEED + 11)
    before = int(capi.lib().tg_tune_get(b"spmm_long_row"))
    out = []
    try:
        for name, n_long, base in (("uniform", 0, 12.0), ("2 rows of 1e6", 2, 10.0)):
            lens = rng.poisson(base, M).clip(1, 64)
            lens[rng.choice(M, n_long, replace=False)] = 1_000_000
            rp = np.r_[0, np.cumsum(lens)].astype(np.int32)
            nnz = int(rp[-1])
            rpd = torch.from_numpy(rp).to(dev)
            cid = torch.from_numpy(rng.integers(0, N, nnz).astype(np.int32)).to(dev)
            vd = torch.from_numpy(rng.normal(size=nnz)).to(dev)
            for split in (1, 0):
                capi.call("tg_tune_set", b"spmm_long_row", 8192 if split else 0)
                plan = SlicePlan(M, N, nnz, rpd, cid)
                for k in (1, 4, 16):
                    X = torch.rand(N, k, dtype=torch.float64, device=dev)
                    Y = torch.empty(M, k, dtype=torch.float64, device=dev)
                    plan.spmm(vd, 1.0, X, k, 0.0, Y, k, k)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    torch.cuda.synchronize()
                    e0.record(s)
                    reps = 5 if (split or n_long == 0) else 1
                    for _ in range(reps):
                        plan.spmm(vd, 1.0, X, k, 0.0, Y, k, k)
                    e1.record(s)
                    torch.cuda.synchronize()
                    out.append({"matrix": name, "nnz": nnz, "split": bool(split), "long_rows": plan.long_rows(),
                                "nrhs": k, "us": round(e0.elapsed_time(e1) * 1e3 / reps, 1)})
                    log(f"long rows: {out[-1]}")
                    del X, Y
                del plan
            del rpd, cid, vd
            torch.cuda.empty_cache()
    finally:
        capi.call("tg_tune_set", b"spmm_long_row", before)
    return {"workload": "1e6 x 1e6, ~12 M nonzeros, uniform columns, FP64, planned SpMM (tg_csr_spmm_plan)",
            "cases": out}


def bsr_section(args, dev, read_gbps):
    """Block-CSR SpMM (KernelMatrixCSR<DenseBlockKernel_Tiny>::Dot, tg_bsr_spmm): R x R FP64 blocks
    on the 3D 7-point grid pattern (block rows g^3), X / Y row-major chunks, alpha = 1, beta = 0.
    Algorithmic bytes nnzb (R R s + 4) + 4 (mb + 1) + mb R k s (X) + mb R k s (Y); HIP events
    on the launch stream around each call."""
    import torch

    from tensors_amd import generators as gen
    from tensors_amd.sparse import KernelMatrixCSR

    from tensors_amd import _capi as capi

    if args.bsr_vector >= 0:
        capi.call("tg_tune_set", b"bsr_vector", args.bsr_vector)
    stream = torch.cuda.current_stream()
    res = []
    for g, R, k in ((96, 3, 1), (96, 3, 4), (64, 4, 4), (48, 8, 4)):
        rp, ci, _ = gen.laplacian_3d(g, 1.0)
        mb = g ** 3
        nnzb = int(rp[-1])
        A = torch.from_numpy(gen.uniform(7007, nnzb * R * R)).to(dev)
        X = torch.from_numpy(gen.uniform(7008, mb * R * k)).to(dev)
        Y = torch.empty(mb * R * k, dtype=torch.float64, device=dev)
        K = KernelMatrixCSR(rp.astype(np.int32), ci, mb, mb, R, R)
        for _ in range(max(1, args.warmup)):
            K.Dot(A, 1.0, X, 0.0, Y, k)
        reps = max(3, args.steps)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            K.Dot(A, 1.0, X, 0.0, Y, k)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        byts = nnzb * (R * R * 8 + 4) + 4 * (mb + 1) + 2 * mb * R * k * 8
        gbps = byts / (ms * 1e-3) / 1e9
        res.append({"grid": g, "block": R, "nrhs": k, "block_rows": mb, "block_nnz": nnzb, "ms": round(ms, 4),
                    "alg_bytes": byts, "GBps": round(gbps, 1), "frac_peak": round(gbps / PEAK_HBM_GBPS, 4),
                    "frac_measured_read": round(gbps / read_gbps, 4)})
        log(f"bsr g={g} R={R} k={k}: {ms:.3f} ms, {gbps:.0f} GB/s")
        del A, X, Y, K
    return {"workload": "R x R FP64 blocks on the 3D 7-point grid pattern, random values, X/Y row-major "
                        "chunks, alpha=1 beta=0 (tg_bsr_spmm)", "cases": res}


def mm_section(args):
    """MatrixMarket write + load (tg_mm_write / tg_mm_read + GPU assembly) of the 3D Laplacian on a
    g^3 grid: MB/s of text and seconds, host-side parsing in parallel."""
    import tempfile

    from tensors_amd import generators as gen
    from tensors_amd.sparse import MatrixCSR

    g = 64
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    A = MatrixCSR(rp, ci, val, g ** 3, g ** 3)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "lap.mtx")
        t0 = time.perf_counter()
        A.write_matrix_market(f)
        tw = time.perf_counter() - t0
        size = os.path.getsize(f)
        t0 = time.perf_counter()
        B = MatrixCSR.from_matrix_market(f)
        tl = time.perf_counter() - t0
    same = bool(np.array_equal(B.Outer(), A.Outer()) and np.array_equal(B.Inner(), A.Inner())
                and np.array_equal(B.Values(), A.Values()))
    return {"workload": f"3D 7-point Laplacian {g}^3 (nnz {int(rp[-1])}), %.17E text", "bytes": size,
            "write_s": round(tw, 3), "load_s": round(tl, 3), "write_MBps": round(size / tw / 1e6, 1),
            "load_MBps": round(size / tl / 1e6, 1), "round_trip_exact": same, "host_threads": os.cpu_count()}


def gmres_section(args, dev):
    """GMRES(30) to 1e-8 (P = I, 100 restarts) on single upwind convection-diffusion systems of the
    serving mix's range (tg_gmres_solve): one warm-up and one timed solve each; ms, Arnoldi steps and
    microseconds per Arnoldi step (restarts x 30 + the last cycle's iterations)."""
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd import generators as gen
    from tensors_amd.batch import BatchGMRES
    from tensors_amd.solvers import GMRES, gmres_solve_mixed
    from tensors_amd.sparse import MatrixCSR

    if args.gmres_delayed >= 0:
        capi.call("tg_tune_set", b"gmres_delayed", args.gmres_delayed)
    res = []
    for g, pe in ((1000, 10.0), (300, 30.0), (100, 3.0)):
        rp, ci, val = gen.convection_diffusion_2d(g, pe)
        n = g * g
        A = MatrixCSR(rp, ci, val, n, n)
        B = torch.from_numpy(gen.uniform(7, n)).to(dev).reshape(n, 1)
        t = 0.0
        for rep_ in range(2):
            X = torch.zeros(n, 1, dtype=torch.float64, device=dev)
            gm = GMRES(n, 30, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ok = gm(A, None, 1.0, B, 0.0, X, 1e-8, 100)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
        steps = gm.restarts * 30 + gm.iter
        # the mixed-precision form (FP32 Arnoldi cycles, FP64 residual / update / stop test)
        tm, okm = 0.0, False
        for rep_ in range(2):
            Xm = torch.zeros(n, 1, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            okm, rsm, stepsm, rrm = gmres_solve_mixed(A, B, Xm, tol=1e-8, restart=30, max_restarts=100)
            torch.cuda.synchronize()
            tm = time.perf_counter() - t0
        res.append({"grid": g, "peclet": pe, "n": n, "succeeded": bool(ok), "restarts": gm.restarts,
                    "arnoldi_steps": steps, "ms": round(t * 1e3, 1), "us_per_step": round(t * 1e6 / max(steps, 1), 1),
                    "mixed": {"succeeded": bool(okm), "restarts": rsm, "fp32_arnoldi_steps": stepsm,
                              "ms": round(tm * 1e3, 1), "true_rel_residual": float(rrm[0]),
                              "speedup_over_fp64": round(t / tm, 3) if tm > 0 else None}})
        log(f"GMRES {res[-1]}")
    # many small systems at once: the batched engine against one tg_gmres_solve after another
    count, gsz = 64, 100
    systems = []
    for i in range(count):
        rp, ci, val = gen.convection_diffusion_2d(gsz, 1.0 + 99.0 * i / (count - 1))
        n = gsz * gsz
        systems.append((torch.from_numpy(rp.astype(np.int32)).to(dev), torch.from_numpy(ci).to(dev),
                        torch.from_numpy(val).to(dev), torch.from_numpy(gen.uniform(70 + i, n)).to(dev).reshape(n, 1),
                        MatrixCSR(rp, ci, val, n, n)))
    Xs = [torch.zeros(gsz * gsz, 1, dtype=torch.float64, device=dev) for _ in systems]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (rpd, cid, vd, B, A), X in zip(systems, Xs):
        GMRES(gsz * gsz, 30, 1)(A, None, 1.0, B, 0.0, X, 1e-8, 100)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    bt = BatchGMRES()
    Xb = [torch.zeros(gsz * gsz, 1, dtype=torch.float64, device=dev) for _ in systems]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ids = [bt.add(rpd, cid, vd, B, X, 1e-8, 30, 100) for (rpd, cid, vd, B, A), X in zip(systems, Xb)]
    bt.run(1)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    okb = sum(bt.result(i)["succeeded"] for i in ids)
    bt.close()
    batched = {"systems": count, "grid": gsz, "peclet": "1..100", "sequential_solves_per_s": round(count / t_seq, 2),
               "batched_solves_per_s": round(count / t_bat, 2), "batched_succeeded": int(okb),
               "speedup": round(t_seq / t_bat, 3)}
    log(f"GMRES batched {batched}")
    return {"workload": "GMRES(30), 1e-8, P = I, 100 restarts, 2D upwind convection-diffusion, 1 RHS "
                        "uniform[-1,1] (tg_gmres_solve; mixed: tg_gmres_solve_mixed)", "cases": res,
            "batched_small_systems": batched}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from tensors_amd import _capi as capi
    from tensors_amd import generators as gen
    from tensors_amd.sparse import SlicePlan

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
    # the path MatrixCSR.Dot takes: tg_csr_spmm_plan (column slices where X is a few L2s)
    if args.slices >= 0:
        capi.call("tg_tune_set", b"spmm_slices", args.slices)
    plan = SlicePlan(M, N, nnz, rp32, cid)
    if not args.no_split:
        # the matrix's values registered with the plan, as MatrixCSR.Dot does: the re-laid
        # 8-slice layout serves the nrhs <= 4 cases with X of 2-64 MB
        plan.set_values(vals[np.float64])
        plan.set_values(vals[np.float32])
    for c in cases:
        c["slices"] = plan.slices(c["Y"].dtype, c["k"], a=c["A"])

    def launch(c):
        return plan.spmm(c["A"], 1.0, c["X"], c["k"], 0.0, c["Y"], c["k"], c["k"])

    # per-case kernel shapes (lanes per row / RHS tile, rows per lane group, column
    # slices) from the project's autotuner, set before each case's launch
    from tensors_amd import autotune

    t_tune = time.perf_counter()
    for c in cases:
        if args.no_tune:
            c["cfg"] = autotune.get_knobs()
        else:
            c["cfg"], _ = autotune.tune_shapes(lambda c=c: launch(c), c["k"], reps=2, keep=False)
    t_tune = time.perf_counter() - t_tune
    if not args.no_tune:
        log(f"SpMM shapes tuned in {t_tune:.1f}s")

    def launch_tuned(c):
        autotune.set_knobs(c["cfg"])
        return launch(c)

    for c in cases:
        autotune.set_knobs(c["cfg"])
        c["slices"] = plan.slices(c["Y"].dtype, c["k"], a=c["A"])
        c["split"] = plan.uses_split(c["A"], c["k"])

    log(f"SpMM suite: nnz={nnz}, {len(cases)} launches per step")
    for _ in range(args.warmup):
        for c in cases:
            launch_tuned(c)
    torch.cuda.synchronize()
    # timed region: events around every launch on the launch stream
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cases]
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        for i, c in enumerate(cases):
            autotune.set_knobs(c["cfg"])
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

    log(f"SpMM suite timed: {value:.1f} GB/s")
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
    for c in cases:
        c["lf_bytes"] = line_fill_bytes(nnz, M, c["k"], 8 if c["dtype"] == "f64" else 4)
    breakdown = [{"dtype": c["dtype"], "nrhs": c["k"], "slices": c["slices"], "relaid": c["split"], "shape": c["cfg"],
                  "ms": round(t, 4),
                  "GBps": round(c["bytes"] / (t * 1e-3) / 1e9, 1),
                  "frac_peak": round(c["bytes"] / (t * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4),
                  "frac_measured_read": round(c["bytes"] / (t * 1e-3) / 1e9 / read_gbps, 4),
                  "line_fill_bytes": c["lf_bytes"],
                  "frac_peak_line_fill": round(c["lf_bytes"] / (t * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)}
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
                               "FP64 and FP32, nrhs 1/4/16/64, alpha=1 beta=0, int32 indices, MatrixCSR.Dot path "
                               "(tg_csr_spmm_plan)",
                   "nnz": nnz, "nrhs": list(KS)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic,
                     "kernel": f"spmm {dc['dtype']} nrhs={dc['k']}",
                     "alg_bytes_per_launch": dc["bytes"],
                     "line_fill_bytes_per_launch": dc["lf_bytes"],
                     "achieved_line_fill": round(dc["lf_bytes"] / (per_case_ms[dom] * 1e-3) / 1e9, 1),
                     "frac_line_fill": round(dc["lf_bytes"] / (per_case_ms[dom] * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)},
        "hbm_measured": {"copy_GBps": round(copy_gbps, 1), "read_GBps": round(read_gbps, 1)},
        "breakdown": breakdown,
    }
    # free the SpMM suite's buffers before the solver section
    del cases, vals, Xs, rp32, cid, plan
    torch.cuda.empty_cache()
    if args.line_rate:
        log("line-rate probe")
        out["line_rate"] = line_rate_section(dev)
    if args.long_rows:
        log("long-row probe")
        out["long_rows"] = long_rows_section(dev)
    if not args.no_bsr:
        log("block-CSR section")
        out["bsr"] = bsr_section(args, dev, read_gbps)
        torch.cuda.empty_cache()
    if not args.no_mm:
        log("MatrixMarket section")
        out["matrix_market"] = mm_section(args)
    if not args.no_cg:
        log("CG section")
        out["cg"] = cg_section(args, dev)
        torch.cuda.empty_cache()
    if not args.no_gmres:
        log("GMRES section")
        out["gmres"] = gmres_section(args, dev)
        torch.cuda.empty_cache()
    if not args.no_batch:
        log("batched CG section")
        out["batch"] = batch_section(args, dev)
        torch.cuda.empty_cache()
    if not args.no_serve:
        log("serving section")
        out["serve"] = serve_section(args, dev)
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(rp, ci, v64, v32, X64h, X32h, args.cpu_reps, args.cpu_warmup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
