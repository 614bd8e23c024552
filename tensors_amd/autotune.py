"""Shape autotuner for the SpMM kernels.

The kernels take their shapes from process-wide knobs (tg_tune_set):
  nrhs == 1: "spmv_kernel" (0 streaming, 1 team-per-row) and "spmv_team" (lanes per
             row, 4..64; the streaming kernel caps it at 32);
  nrhs >= 2: "spmm_group" (lanes per RHS tile, 1..64; columns beyond G*V are done in
             further tiles) and "spmm_rows" (1: the wave's 64/G lane groups split one
             row's nonzeros; 2: each lane group owns its own row);
0 means "pick from the matrix" (mean row length, nrhs).  tune_spmm times every
candidate on the caller's own matrix and block with HIP events on the launch stream,
keeps the fastest, and leaves it set (or restores the previous setting).  Results
are identical for every shape up to floating-point summation order within a row's
lane groups.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi
from .sparse import MatrixCSR

KNOBS = ("spmv_kernel", "spmv_team", "spmm_group", "spmm_rows", "spmm_slices", "spmm_split")


def get_knobs():
    return {k: int(capi.lib().tg_tune_get(k.encode())) for k in KNOBS}


def set_knobs(cfg):
    for k, v in cfg.items():
        capi.call("tg_tune_set", k.encode(), int(v))


def candidates(nrhs):
    if nrhs == 1:
        out = [{"spmv_kernel": 0, "spmv_team": 0}, {"spmv_kernel": 1, "spmv_team": 0}]
        out += [{"spmv_kernel": 0, "spmv_team": t} for t in (1, 2, 4, 8, 16, 32)]
        out += [{"spmv_kernel": 1, "spmv_team": t} for t in (4, 8, 16, 32, 64)]
        return out
    gmax = 1
    while gmax < nrhs and gmax < 64:
        gmax *= 2
    out = [{"spmm_group": 0, "spmm_rows": 0}]
    for rows in (1, 2):  # row per wave / row per lane group
        out += [{"spmm_group": g, "spmm_rows": rows} for g in (1, 2, 4, 8, 16, 32, 64)
                if g <= gmax and not (rows == 2 and g == 64)]
    return out


def slice_candidates(nrhs):
    """Column-slice choices of the planned path (nrhs <= 4 only): none, the 2/4/8-slice
    cut of the original layout, and the re-laid 8-slice layout (used where values are
    registered with the plan and X is 2-64 MB)."""
    if nrhs > 4:
        return [{"spmm_slices": 1, "spmm_split": 0}]
    return [{"spmm_slices": s, "spmm_split": 0} for s in (1, 2, 4, 8)] + [{"spmm_slices": 0, "spmm_split": 1}]


def tune_shapes(launch, nrhs, reps=5, keep=True):
    """Times launch() -- one product enqueued on the current stream -- for every kernel
    shape with column slicing off, then every slice count with the best shape; returns
    (best_cfg, table of (cfg, ms)) and leaves the best set (keep) or restores the
    previous knobs."""
    stream = torch.cuda.current_stream()
    before = get_knobs()
    table = []

    def timed(cfg):
        set_knobs(cfg)
        launch()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(reps):
            launch()
        ev[1].record(stream)
        torch.cuda.synchronize()
        table.append((dict(cfg), ev[0].elapsed_time(ev[1]) / reps))
        return table[-1][1]

    try:
        shapes = [dict(c, spmm_slices=1, spmm_split=0) for c in candidates(nrhs)]
        best = min(shapes, key=timed)
        for sc in slice_candidates(nrhs)[1:]:
            timed(dict(best, **sc))
    finally:
        set_knobs(before)
    best = min(table, key=lambda t: t[1])[0]
    if keep:
        set_knobs(best)
    return best, table


def tune_spmm(A: MatrixCSR, X: torch.Tensor, Y: torch.Tensor, reps: int = 5, keep: bool = True):
    """Times Y <- A X (the MatrixCSR.Dot path, tg_csr_spmm_plan) for every kernel shape with
    column slicing off, then every slice count with the best shape; returns (best_cfg,
    table of (cfg, ms)).  X, Y: CUDA tensors of A's value dtype, n x nrhs / m x nrhs (or
    1-D for nrhs = 1)."""
    nrhs = 1 if X.dim() == 1 else X.shape[1]
    _, _, vals = A.device_arrays(X.dtype)
    plan = A.Plan()
    ld_x = 1 if X.dim() == 1 else X.stride(0)
    ld_y = 1 if Y.dim() == 1 else Y.stride(0)
    return tune_shapes(lambda: plan.spmm(vals, 1.0, X, ld_x, 0.0, Y, ld_y, nrhs), nrhs, reps, keep)


BATCH_KNOBS = ("batch_tile_rows", "batch_spmm")


def get_batch_knobs():
    return {k: int(capi.lib().tg_tune_get(k.encode())) for k in BATCH_KNOBS}


def batch_candidates(tile_rows=(256, 512, 1024, 2048), spmm_forms=(2, 1, 0)):
    """Rows per workgroup tile x SpMM kernel of the batched CG engine.  Every SpMM form
    gives bitwise the same iterates; the tile height changes the order in which a
    system's tile partials are added (rounding only)."""
    return [{"batch_tile_rows": t, "batch_spmm": f} for t in tile_rows for f in spmm_forms]


def tune_batch(systems, iters=24, warm=4, mixed=1, candidates=None, keep=True):
    """Times the batched CG engine on the caller's systems for every candidate shape and
    keeps the fastest.  systems: list of (rp, ci, val, diag, B) CUDA tensors (int32 rp/ci,
    float64 values / diagonal / right-hand sides); each candidate runs a fresh batch that
    holds all of them for `warm` untimed and `iters` timed iterations (tol = 0, so no
    system leaves early) in mode `mixed` (0 FP64, 1 mixed).  Returns (best_cfg, table of
    (cfg, ms per iteration)); the knobs apply to systems added after the call."""
    import time

    from .batch import BatchCG

    cands = candidates if candidates is not None else batch_candidates()
    before = get_batch_knobs()
    table = []
    try:
        for cfg in cands:
            set_knobs(cfg)
            bt = BatchCG()
            Xs = []
            for rp, ci, val, d, B in systems:
                X = torch.zeros_like(B)
                Xs.append(X)
                bt.add(rp, ci, val, B, X, 0.0, warm + iters + 1, diag=d, mixed=mixed == 1)
            bt.step(warm)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bt.step(iters)  # synchronises the batch's stream
            table.append((dict(cfg), (time.perf_counter() - t0) * 1e3 / iters))
            bt.close()
            del Xs
    finally:
        set_knobs(before)
    best = min(table, key=lambda t: t[1])[0]
    if keep:
        set_knobs(best)
    return best, table
