"""Python mirror of the ConjugateGradient / GMRES drivers over the C ABI.

``ConjugateGradient(n, max_iter, nrhs)`` keeps its device workspace between calls, as the
reference keeps its Tensor2 workspaces (ConjugateGradient.hpp:38-52, 64-88); calling it
runs tg_cg_solve on CUDA tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi
from .sparse import MatrixCSR, stream_handle

_TG = {torch.float64: capi.F64, torch.float32: capi.F32}


class ConjugateGradient:
    """ConjugateGradient<VarSize,Scal,Int>(n, max_iter, nrhs) (ConjugateGradient.hpp:64-88)."""

    def __init__(self, n, max_iter, nrhs=1, dtype=torch.float64, device="cuda"):
        self.n, self.max_iter, self.nrhs = int(n), int(max(0, max_iter)), int(nrhs)
        self.dtype, self.device = dtype, device
        self._ws = None
        self.iter = 0
        self.succeeded = False
        self.rel_res = np.zeros(self.nrhs)

    def workspace(self, prec_kind):
        nb = capi.lib().tg_cg_workspace_bytes(self.n, self.nrhs, _TG[self.dtype], prec_kind)
        if self._ws is None or self._ws.numel() < nb:
            self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        return self._ws, nb

    def __call__(self, A: MatrixCSR, P, a, B, b, X, tol, use_initial_guess=False, stream=None):
        """X <- a A^{-1} B + b X.  P: None (identity) or a CUDA tensor d (z = d * r)."""
        prec_kind = 0 if P is None else 1
        ws, nb = self.workspace(prec_kind)
        rp, ci, vals = A.device_arrays(self.dtype)
        ok = C.c_int32(0)
        it = C.c_int32(0)
        rr = np.zeros(self.nrhs, np.float64)
        capi.call("tg_cg_solve", self.n, A.NonzeroCount(), rp.data_ptr(),
                  capi.I64 if rp.dtype == torch.int64 else capi.I32, ci.data_ptr(),
                  vals.data_ptr() if vals is not None else None, _TG[self.dtype], prec_kind,
                  P.data_ptr() if P is not None else None, self.nrhs, self.max_iter, float(a), B.data_ptr(),
                  B.stride(0) if B.dim() == 2 else 1, float(b), X.data_ptr(),
                  X.stride(0) if X.dim() == 2 else 1, float(tol), int(use_initial_guess), ws.data_ptr(), nb,
                  C.byref(ok), C.byref(it), rr.ctypes.data, stream_handle(stream))
        self.iter, self.succeeded, self.rel_res = it.value, bool(ok.value), rr
        return self.succeeded

    def IterationCount(self):
        return self.iter

    def RelativeResiduals(self):
        return self.rel_res

    def Residuals(self):
        """sqrt|rho_k| of the last solve (ConjugateGradient.hpp:447-456), via tg_cg_residuals."""
        out = np.zeros(self.nrhs)
        capi.call("tg_cg_residuals", self._ws.data_ptr(), self._ws.numel(), self.nrhs, out.ctypes.data, None, None)
        return out


class GMRES:
    """GMRES<VarSize,Scal,Int,Side>(n, max_iter, nrhs) (GMRES.hpp:82-108); side 0 = Left, 1 = Right."""

    def __init__(self, n, max_iter, nrhs=1, side=0, dtype=torch.float64, device="cuda"):
        self.n, self.max_iter, self.nrhs, self.side = int(n), int(max_iter), int(nrhs), int(side)
        self.dtype, self.device = dtype, device
        self._ws = None
        self.iter = 0
        self.restarts = 0
        self.succeeded = False
        self.rel_res = np.zeros(self.nrhs)

    def workspace(self):
        nb = capi.lib().tg_gmres_workspace_bytes(self.n, self.max_iter, self.nrhs, _TG[self.dtype])
        if self._ws is None or self._ws.numel() < nb:
            self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        return self._ws, nb

    def __call__(self, A: MatrixCSR, P, a, B, b, X, tol, max_restarts, use_initial_guess=False, stream=None):
        prec_kind = 0 if P is None else 1
        ws, nb = self.workspace()
        rp, ci, vals = A.device_arrays(self.dtype)
        ok = C.c_int32(0)
        it = C.c_int32(0)
        rs = C.c_int32(0)
        rr = np.zeros(self.nrhs, np.float64)
        capi.call("tg_gmres_solve", self.n, A.NonzeroCount(), rp.data_ptr(),
                  capi.I64 if rp.dtype == torch.int64 else capi.I32, ci.data_ptr(),
                  vals.data_ptr() if vals is not None else None, _TG[self.dtype], prec_kind,
                  P.data_ptr() if P is not None else None, self.side, self.nrhs, self.max_iter, float(a),
                  B.data_ptr(), B.stride(0) if B.dim() == 2 else 1, float(b), X.data_ptr(),
                  X.stride(0) if X.dim() == 2 else 1, float(tol), int(max_restarts), int(use_initial_guess),
                  ws.data_ptr(), nb, C.byref(ok), C.byref(it), C.byref(rs), rr.ctypes.data,
                  stream_handle(stream))
        self.iter, self.restarts, self.succeeded, self.rel_res = it.value, rs.value, bool(ok.value), rr
        return self.succeeded

    def IterationCount(self):
        return self.iter

    def RestartCount(self):
        return self.restarts

    def RelativeResiduals(self):
        return self.rel_res

    def _parts(self):
        """Device views (Q, H, beta, solution) of the workspace (tg_gmres_workspace_parts)."""
        q, h, bt, x = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        capi.call("tg_gmres_workspace_parts", self._ws.data_ptr(), self._ws.numel(), self.n, self.max_iter,
                  self.nrhs, _TG[self.dtype], C.byref(q), C.byref(h), C.byref(bt), C.byref(x))
        m = max(1, min(self.max_iter, self.n + 1))
        vs = 8 if self.dtype == torch.float64 else 4
        base = self._ws.data_ptr()

        def view(ptr, shape, dtype, esize):
            off = ptr.value - base
            count = int(np.prod(shape))
            return self._ws[off:off + count * esize].view(dtype).view(*shape)
        return (view(q, (m + 1, self.n, self.nrhs), self.dtype, vs),
                view(h, (m + 1, m, self.nrhs), torch.float64, 8),
                view(bt, (m + 1, self.nrhs), torch.float64, 8),
                view(x, (self.n, self.nrhs), self.dtype, vs))

    def GetBasis(self):
        """Basis Q of the last cycle, (m+1) x n x nrhs (GMRES.hpp:820-823)."""
        return self._parts()[0].clone()

    def GetHessenbergMatrix(self):
        """H after the Givens rotations, (m+1) x m x nrhs (GMRES.hpp:825-828)."""
        return self._parts()[1].clone()

    def Solution(self):
        """Solution before the a/b combination, n x nrhs (GMRES.hpp:830-833)."""
        return self._parts()[3].clone()

    def Residuals(self):
        """|beta[iter][k]| (GMRES.hpp:779-789)."""
        return self._parts()[2][self.iter].abs().cpu().numpy()


def csr_callback(A: MatrixCSR, dtype=torch.float64):
    """An operator callback (tg_apply_fn) that applies the CSR matrix A with tg_csr_spmm:
    the analogue of the reference's `[&](cptr x, mptr y){ A.Dot(1, x, 0, y, nrhs); }`
    lambdas (Mathematica/Dev_CG.nb:104-123).  Keep the returned object alive."""
    rp, ci, vals = A.device_arrays(dtype)
    tg = _TG[dtype]
    rpt = capi.I64 if rp.dtype == torch.int64 else capi.I32

    def fn(ctx, x, y, k, stream):
        return capi.lib().tg_csr_spmm(A.m, A.n, A.NonzeroCount(), rp.data_ptr(), rpt, ci.data_ptr(),
                                      vals.data_ptr() if vals is not None else None, tg, 1.0, x, k, 0.0, y, k, k,
                                      0, None, 0, stream)

    return capi.APPLY_FN(fn)


def cg_solve_op(n, nrhs, max_iter, apply_A, apply_P, a, B, b, X, tol, use_initial_guess=False,
                dtype=torch.float64, stream=None):
    """tg_cg_solve_op with callbacks; returns (succeeded, iterations, rel_res)."""
    nb = capi.lib().tg_cg_workspace_bytes(n, nrhs, _TG[dtype], 1 if apply_P is not None else 0)
    ws = torch.empty(nb, dtype=torch.uint8, device=B.device)
    ok, it = C.c_int32(0), C.c_int32(0)
    rr = np.zeros(nrhs)
    capi.call("tg_cg_solve_op", n, apply_A, None, apply_P, None, _TG[dtype], nrhs, max_iter, float(a),
              B.data_ptr(), B.stride(0), float(b), X.data_ptr(), X.stride(0), float(tol), int(use_initial_guess),
              ws.data_ptr(), nb, C.byref(ok), C.byref(it), rr.ctypes.data, stream_handle(stream))
    return bool(ok.value), it.value, rr


def gmres_solve_op(n, nrhs, max_iter, apply_A, apply_P, a, B, b, X, tol, max_restarts, side=0,
                   use_initial_guess=False, dtype=torch.float64, stream=None):
    """tg_gmres_solve_op with callbacks; returns (succeeded, iterations, restarts, rel_res)."""
    nb = capi.lib().tg_gmres_workspace_bytes(n, max_iter, nrhs, _TG[dtype])
    ws = torch.empty(nb, dtype=torch.uint8, device=B.device)
    ok, it, rs = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    rr = np.zeros(nrhs)
    capi.call("tg_gmres_solve_op", n, apply_A, None, apply_P, None, _TG[dtype], side, nrhs, max_iter, float(a),
              B.data_ptr(), B.stride(0), float(b), X.data_ptr(), X.stride(0), float(tol), int(max_restarts),
              int(use_initial_guess), ws.data_ptr(), nb, C.byref(ok), C.byref(it), C.byref(rs), rr.ctypes.data,
              stream_handle(stream))
    return bool(ok.value), it.value, rs.value, rr


def cg_solve_mixed(A: MatrixCSR, B, X, tol=1e-10, max_outer=100, max_inner=20000, replace_drop=0.1, a=1.0, b=0.0,
                   P=None, use_initial_guess=False, a32=None, stream=None):
    """tg_cg_solve_mixed: FP32 CG with FP64 residual replacement to ||B - A X|| / ||B|| <= tol.
    Returns (succeeded, replacements, fp32_iterations, rel_res).  a32: optional FP32 values (CUDA)."""
    n, k = X.shape[0], (X.shape[1] if X.dim() == 2 else 1)
    rp, ci, vals = A.device_arrays(torch.float64)
    if a32 is None:
        _, _, a32 = A.device_arrays(torch.float32)
    nb = capi.lib().tg_cg_mixed_workspace_bytes(n, A.NonzeroCount(), k, 0 if P is None else 1, 1)
    ws = torch.empty(nb, dtype=torch.uint8, device=X.device)
    ok, outer, inner = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    rr = np.zeros(k)
    capi.call("tg_cg_solve_mixed", n, A.NonzeroCount(), rp.data_ptr(),
              capi.I64 if rp.dtype == torch.int64 else capi.I32, ci.data_ptr(), vals.data_ptr(), a32.data_ptr(),
              0 if P is None else 1, P.data_ptr() if P is not None else None, k, max_outer, max_inner,
              float(replace_drop), float(a), B.data_ptr(), B.stride(0) if B.dim() == 2 else 1, float(b),
              X.data_ptr(), X.stride(0) if X.dim() == 2 else 1, float(tol), int(use_initial_guess),
              ws.data_ptr(), nb, C.byref(ok), C.byref(outer), C.byref(inner), rr.ctypes.data,
              stream_handle(stream))
    return bool(ok.value), outer.value, inner.value, rr


def gmres_solve_mixed(A: MatrixCSR, B, X, tol=1e-8, restart=30, max_restarts=100, a=1.0, b=0.0, P=None, side=0,
                      use_initial_guess=False, a32=None, stream=None):
    """tg_gmres_solve_mixed: GMRES(restart) cycles in FP32 with the residual, solution and
    stop test in FP64 at every restart.  Returns (succeeded, restarts, fp32_steps, rel_res).
    P: None or a CUDA float64 diagonal (z = d * r); a32: optional FP32 values (CUDA)."""
    n, k = X.shape[0], (X.shape[1] if X.dim() == 2 else 1)
    rp, ci, vals = A.device_arrays(torch.float64)
    if a32 is None:
        _, _, a32 = A.device_arrays(torch.float32)
    nb = capi.lib().tg_gmres_mixed_workspace_bytes(n, A.NonzeroCount(), restart, k, 0 if P is None else 1, 1)
    ws = torch.empty(nb, dtype=torch.uint8, device=X.device)
    ok, rs, inner = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    rr = np.zeros(k)
    capi.call("tg_gmres_solve_mixed", n, A.NonzeroCount(), rp.data_ptr(),
              capi.I64 if rp.dtype == torch.int64 else capi.I32, ci.data_ptr(), vals.data_ptr(), a32.data_ptr(),
              0 if P is None else 1, P.data_ptr() if P is not None else None, int(side), k, int(restart),
              int(max_restarts), float(a), B.data_ptr(), B.stride(0) if B.dim() == 2 else 1, float(b),
              X.data_ptr(), X.stride(0) if X.dim() == 2 else 1, float(tol), int(use_initial_guess),
              ws.data_ptr(), nb, C.byref(ok), C.byref(rs), C.byref(inner), rr.ctypes.data, stream_handle(stream))
    return bool(ok.value), rs.value, inner.value, rr
