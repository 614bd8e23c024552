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

KNOBS = ("spmv_kernel", "spmv_team", "spmm_group", "spmm_rows", "spmm_slices")


def get_knobs():
    return {k: int(capi.lib().tg_tune_get(k.encode())) for k in KNOBS}


def set_knobs(cfg):
    for k, v in cfg.items():
        capi.call("tg_tune_set", k.encode(), int(v))


def candidates(nrhs):
    if nrhs == 1:
        out = [{"spmv_kernel": 0, "spmv_team": 0}, {"spmv_kernel": 1, "spmv_team": 0}]
        out += [{"spmv_kernel": 0, "spmv_team": t} for t in (4, 8, 16, 32)]
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
    """Column-slice counts of the planned path (nrhs <= 4 only)."""
    return [{"spmm_slices": s} for s in ((1, 2, 4, 8) if nrhs <= 4 else (1,))]


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
    stream = torch.cuda.current_stream()
    before = get_knobs()
    table = []

    def timed(cfg):
        set_knobs(cfg)
        plan.spmm(vals, 1.0, X, ld_x, 0.0, Y, ld_y, nrhs)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(reps):
            plan.spmm(vals, 1.0, X, ld_x, 0.0, Y, ld_y, nrhs)
        ev[1].record(stream)
        torch.cuda.synchronize()
        table.append((dict(cfg), ev[0].elapsed_time(ev[1]) / reps))
        return table[-1][1]

    try:
        shapes = [dict(c, spmm_slices=1) for c in candidates(nrhs)]
        best = min(shapes, key=timed)
        for sc in slice_candidates(nrhs)[1:]:
            timed(dict(best, **sc))
    finally:
        set_knobs(before)
    best = min(table, key=lambda t: t[1])[0]
    if keep:
        set_knobs(best)
    return best, table
