"""Synthetic workloads (host numpy arrays) from the library's counter-based
generators (tg_gen_*, tensors_amd/csrc/generators.cpp)."""
from __future__ import annotations

import math

import numpy as np

from . import _capi as capi


def _p(a):
    return a.ctypes.data if a is not None else None


def uniform(seed: int, count: int, dtype=np.float64, offset: int = 0) -> np.ndarray:
    """Uniform[-1,1] values, counter-based (same output for any thread count)."""
    out = np.empty(count, dtype=dtype)
    capi.call("tg_gen_uniform", seed, offset, count, capi.F64 if dtype == np.float64 else capi.F32, _p(out))
    return out


def random_csr(m: int, n: int, mean: float = 32.0, lo: int = 1, hi: int = 512, seed: int = 2002,
               want64: bool = True, want32: bool = False):
    """Random CSR: Poisson(mean) row lengths clipped to [lo, hi], distinct sorted
    columns per row, N(0,1) values.  Returns (rp int64, ci int32, val64, val32)."""
    rp = np.empty(m + 1, dtype=np.int64)
    nnz = capi.call("tg_gen_random_csr_rowptr", m, mean, lo, hi, seed, _p(rp))
    ci = np.empty(nnz, dtype=np.int32)
    v64 = np.empty(nnz, dtype=np.float64) if want64 else None
    v32 = np.empty(nnz, dtype=np.float32) if want32 else None
    capi.call("tg_gen_random_csr_fill", m, n, _p(rp), seed, _p(ci), _p(v64), _p(v32))
    return rp, ci, v64, v32


def laplacian_3d(g: int, shift: float = 1e-2):
    """7-point graph Laplacian + shift on a g^3 grid: (rp int64, ci int32, val float64)."""
    N = g ** 3
    nnz = 7 * N - 6 * g * g
    rp = np.empty(N + 1, dtype=np.int64)
    ci = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    got = capi.call("tg_gen_laplacian_3d", g, shift, _p(rp), _p(ci), _p(val))
    assert got == nnz
    return rp, ci, val


def weighted_grid(dims: int, sides, shift: float = 1e-3, seed: int = 4004):
    """Weighted 2D/3D grid Laplacian with exp(N(0,1)) edge weights + shift (SPD)."""
    s = list(sides) + [1] * (3 - len(sides))
    N = s[0] * s[1] * (s[2] if dims == 3 else 1)
    cap = (2 * dims + 1) * N
    rp = np.empty(N + 1, dtype=np.int64)
    ci = np.empty(cap, dtype=np.int32)
    val = np.empty(cap, dtype=np.float64)
    nnz = capi.call("tg_gen_weighted_grid", dims, s[0], s[1], s[2], shift, seed, _p(rp), _p(ci), _p(val))
    return rp, ci[:nnz].copy(), val[:nnz].copy()


def convection_diffusion_2d(g: int, peclet: float, theta: float = math.pi / 6):
    """Upwind convection-diffusion on a g x g grid (nonsymmetric M-matrix)."""
    N = g * g
    nnz = 5 * N - 4 * g
    rp = np.empty(N + 1, dtype=np.int64)
    ci = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    got = capi.call("tg_gen_convection_diffusion_2d", g, peclet, theta, _p(rp), _p(ci), _p(val))
    assert got == nnz
    return rp, ci, val
