"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test-side checker only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

i32, i64, dbl, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p

_SIG = {
    "orc_job_pointer": (i64, [i64, i64, i64]),
    "orc_job_pointers_costs": (None, [i64, vp, i64, vp]),
    "orc_multiply_dense_f64": (C.c_int, [vp, vp, vp, i32, i32, dbl, vp, i64, dbl, vp, i64, i32, i32, i32]),
    "orc_multiply_dense_f32": (C.c_int, [vp, vp, vp, i32, i32, C.c_float, vp, i64, C.c_float, vp, i64, i32,
                                         i32, i32]),
    "orc_bsr_spmm_f64": (C.c_int, [i64, vp, vp, i32, i32, vp, i32, dbl, vp, i32, dbl, vp, i32, i32]),
    "orc_bsr_spmm_f32": (C.c_int, [i64, vp, vp, i32, i32, vp, i32, C.c_float, vp, i32, C.c_float, vp, i32, i32]),
    "orc_bsr_fill_lower": (None, [i64, vp, vp, i32, vp]),
    "orc_from_triples": (i64, [i64, vp, vp, vp, i32, i64, i64, C.c_int, C.c_int, vp, vp, vp, i64]),
    "orc_grid_laplacian": (i64, [i32, dbl, i32, vp, vp, vp, i64]),
    "orc_transpose": (C.c_int, [i64, i64, vp, vp, vp, i32, vp, vp, vp]),
    "orc_spgemm_expanded_count": (i64, [i64, vp, vp, vp]),
    "orc_spgemm": (i64, [i64, i64, i64, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, i64]),
    "orc_trsv_upper": (None, [i32, vp, i32, vp]),
    "orc_cg": (C.c_int, [i32, vp, vp, vp, i32, vp, i32, i32, i32, dbl, vp, i64, dbl, vp, i64, dbl, i32, vp,
                         vp]),
    "orc_gmres": (C.c_int, [i32, vp, vp, vp, i32, vp, i32, i32, i32, i32, dbl, vp, i64, dbl, vp, i64, dbl, i32,
                            i32, vp, vp, vp]),
    "orc_set_pinning": (None, [i32]),
    "orc_cg_ex": (C.c_int, [i32, vp, vp, vp, i32, vp, i32, i32, i32, dbl, vp, i64, dbl, vp, i64, dbl, i32, vp,
                            vp, vp]),
    "orc_gmres_ex": (C.c_int, [i32, vp, vp, vp, i32, vp, i32, i32, i32, i32, dbl, vp, i64, dbl, vp, i64, dbl, i32,
                               i32, vp, vp, vp, vp, vp, vp, vp]),
}

_libs = {}


def lib(fast: bool = False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(ORACLE_DIR, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        for fn, (res, args) in _SIG.items():
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        _libs[name] = L
    return _libs[name]


def p(a):
    return None if a is None else a.ctypes.data


def spmm(rp, ci, a, m, n, alpha, X, ldX, beta, Y, ldY, nrhs, NRHS=0, threads=4, fast=False):
    """In-place oracle SpMM on numpy arrays (rp int64, ci int32). Returns the status."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    if Y.dtype == np.float64:
        return lib(fast).orc_multiply_dense_f64(p(rp), p(ci), p(a), m, n, alpha, p(X), ldX, beta, p(Y), ldY,
                                                nrhs, NRHS, threads)
    return lib(fast).orc_multiply_dense_f32(p(rp), p(ci), p(a), m, n, alpha, p(X), ldX, beta, p(Y), ldY, nrhs,
                                            NRHS, threads)


def bsr_spmm(rp, ci, R, C_, A, alpha, X, beta, Y, nrhs, a_rm=True, x_rm=True, y_rm=True):
    """Oracle KernelMatrixCSR<DenseBlockKernel>::Dot; Y (contiguous, A's dtype) updated in place."""
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    mb = len(rp) - 1
    assert A.flags.c_contiguous and X.flags.c_contiguous and Y.flags.c_contiguous
    f = lib().orc_bsr_spmm_f64 if Y.dtype == np.float64 else lib().orc_bsr_spmm_f32
    return f(mb, p(rp), p(ci), R, C_, p(A), int(a_rm), alpha, p(X), int(x_rm), beta, p(Y), int(y_rm), nrhs)


def bsr_fill_lower(rp, ci, R, A):
    """Oracle FillLowerTriangleFromUpperTriangle on float64 block values (in place)."""
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    assert A.dtype == np.float64 and A.flags.c_contiguous
    lib().orc_bsr_fill_lower(len(rp) - 1, p(rp), p(ci), R, p(A))


def grid_laplacian(g, mass, parallel_threads=0):
    cap = 5 * g * g
    rp = np.empty(g * g + 1, np.int64)
    ci = np.empty(cap, np.int32)
    val = np.empty(cap, np.float64)
    nnz = lib().orc_grid_laplacian(g, mass, parallel_threads, p(rp), p(ci), p(val), cap)
    assert nnz >= 0
    return rp, ci[:nnz].copy(), val[:nnz].copy()


def from_triples(I, J, V, m, n, lists=1, compress=True, symmetrize=False):
    I = np.ascontiguousarray(I, np.int32)
    J = np.ascontiguousarray(J, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    cap = max(1, len(I) * (2 if symmetrize else 1))
    rp = np.empty(m + 1, np.int64)
    ci = np.empty(cap, np.int32)
    val = np.empty(cap, np.float64)
    nnz = lib().orc_from_triples(len(I), p(I), p(J), p(V), lists, m, n, int(compress), int(symmetrize), p(rp),
                                 p(ci), p(val), cap)
    assert nnz >= 0
    return rp, ci[:nnz].copy(), val[:nnz].copy()


def transpose(m, n, rp, ci, val, threads=3):
    rp = np.ascontiguousarray(rp, np.int64)
    nnz = int(rp[m])
    orp = np.empty(n + 1, np.int64)
    oci = np.empty(max(nnz, 1), np.int32)
    oval = np.empty(max(nnz, 1), np.float64) if val is not None else None
    lib().orc_transpose(m, n, p(rp), p(ci), p(val), threads, p(orp), p(oci), p(oval))
    return orp, oci[:nnz].copy(), (oval[:nnz].copy() if val is not None else None)


def spgemm(m, k, n, A, B, threads=3, values=True):
    Arp, Aci, Aval = A
    Brp, Bci, Bval = B
    Arp = np.ascontiguousarray(Arp, np.int64)
    Brp = np.ascontiguousarray(Brp, np.int64)
    cap = max(1, lib().orc_spgemm_expanded_count(m, p(Arp), p(Aci), p(Brp)))
    Crp = np.empty(m + 1, np.int64)
    Cci = np.empty(cap, np.int32)
    Cval = np.empty(cap, np.float64)
    nnz = lib().orc_spgemm(m, k, n, p(Arp), p(Aci), p(Aval) if values else None, p(Brp), p(Bci),
                           p(Bval) if values else None, threads, p(Crp), p(Cci), p(Cval), cap)
    assert nnz >= 0
    return Crp, Cci[:nnz].copy(), (Cval[:nnz].copy() if values else None)


def _solver_arrays(rp, ci, val, B, prec_diag):
    """The solvers' C signatures take int64 rp, int32 ci and float64 values / B / diagonal."""
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    val = None if val is None else np.ascontiguousarray(val, np.float64)
    B = np.asarray(B, np.float64)
    if B.ndim == 2 and B.strides[1] != 8:
        B = np.ascontiguousarray(B)
    prec_diag = None if prec_diag is None else np.ascontiguousarray(prec_diag, np.float64)
    return rp, ci, val, B, prec_diag


def cg(n, rp, ci, val, B, X, nrhs, max_iter, tol, a=1.0, b=0.0, use_guess=False, prec_diag=None, threads=4,
       fast=False):
    """Oracle CG; X (n x nrhs, C order, float64) updated in place.  Returns (succeeded, iters,
    rel_res).  rp, ci, val, B and prec_diag are converted to the C types (int64, int32,
    float64) first."""
    rp, ci, val, B, prec_diag = _solver_arrays(rp, ci, val, B, prec_diag)
    assert X.dtype == np.float64 and X.strides[1] == 8
    it = np.zeros(1, np.int32)
    rr = np.zeros(nrhs, np.float64)
    ok = lib(fast).orc_cg(n, p(rp), p(ci), p(val), 1 if prec_diag is not None else 0, p(prec_diag), nrhs,
                          max_iter, threads, a, p(B), B.strides[0] // 8, b, p(X), X.strides[0] // 8, tol,
                          int(use_guess), p(it), p(rr))
    return ok, int(it[0]), rr


def gmres(n, rp, ci, val, B, X, nrhs, max_iter, tol, max_restarts, a=1.0, b=0.0, use_guess=False,
          prec_diag=None, side=0, threads=4, fast=False):
    """Oracle GMRES; X updated in place.  Returns (succeeded, iters, restarts, rel_res)."""
    rp, ci, val, B, prec_diag = _solver_arrays(rp, ci, val, B, prec_diag)
    assert X.dtype == np.float64 and X.strides[1] == 8
    it = np.zeros(1, np.int32)
    rs = np.zeros(1, np.int32)
    rr = np.zeros(nrhs, np.float64)
    ok = lib(fast).orc_gmres(n, p(rp), p(ci), p(val), 1 if prec_diag is not None else 0, p(prec_diag), side,
                             nrhs, max_iter, threads, a, p(B), B.strides[0] // 8, b, p(X), X.strides[0] // 8,
                             tol, max_restarts, int(use_guess), p(it), p(rs), p(rr))
    return ok, int(it[0]), int(rs[0]), rr


def cg_ex(n, rp, ci, val, B, X, nrhs, max_iter, tol, a=1.0, b=0.0, use_guess=False, prec_diag=None, threads=4):
    """oracle.cg plus ConjugateGradient::Residuals: returns (succeeded, iters, rel_res, abs_res)."""
    rp, ci, val, B, prec_diag = _solver_arrays(rp, ci, val, B, prec_diag)
    assert X.dtype == np.float64 and X.strides[1] == 8
    it = np.zeros(1, np.int32)
    rr = np.zeros(nrhs, np.float64)
    ar = np.zeros(nrhs, np.float64)
    ok = lib().orc_cg_ex(n, p(rp), p(ci), p(val), 1 if prec_diag is not None else 0, p(prec_diag), nrhs,
                         max_iter, threads, a, p(B), B.strides[0] // 8, b, p(X), X.strides[0] // 8, tol,
                         int(use_guess), p(it), p(rr), p(ar))
    return ok, int(it[0]), rr, ar


def gmres_ex(n, rp, ci, val, B, X, nrhs, max_iter, tol, max_restarts, a=1.0, b=0.0, use_guess=False,
             prec_diag=None, side=0, threads=4):
    """oracle.gmres plus the final-cycle state of GetBasis / GetHessenbergMatrix / Solution / beta.
    Returns (succeeded, iters, restarts, rel_res, Q, H, beta, solution) with m = min(max_iter, n+1)."""
    rp, ci, val, B, prec_diag = _solver_arrays(rp, ci, val, B, prec_diag)
    assert X.dtype == np.float64 and X.strides[1] == 8
    m = min(max_iter, n + 1)
    it = np.zeros(1, np.int32)
    rs = np.zeros(1, np.int32)
    rr = np.zeros(nrhs, np.float64)
    Q = np.zeros((m + 1, n, nrhs))
    H = np.zeros((m + 1, m, nrhs))
    beta = np.zeros((m + 1, nrhs))
    sol = np.zeros((n, nrhs))
    ok = lib().orc_gmres_ex(n, p(rp), p(ci), p(val), 1 if prec_diag is not None else 0, p(prec_diag), side,
                            nrhs, max_iter, threads, a, p(B), B.strides[0] // 8, b, p(X), X.strides[0] // 8,
                            tol, max_restarts, int(use_guess), p(it), p(rs), p(rr), p(Q), p(H), p(beta), p(sol))
    return ok, int(it[0]), int(rs[0]), rr, Q, H, beta, sol
