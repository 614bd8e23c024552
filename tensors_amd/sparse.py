"""Python mirror of ``Tensors::Sparse::MatrixCSR`` over the gfx950 C ABI.

Host arrays stay authoritative (as in the reference); a device mirror of the row
pointers, column indices and values is made on first use and reused, like the lazily
built job pointers of ``PatternCSR::RequireJobPtr`` (src/Sparse/PatternCSR.hpp:389-398).
Dense blocks passed to ``Dot`` are torch CUDA tensors (row-major, a leading dimension
given by their row stride).  torch only supplies device memory and streams here.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import torch

from . import _capi as capi

_TORCH_TO_TG = {torch.float64: capi.F64, torch.float32: capi.F32}


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _eprint(msg: str) -> None:
    print(msg, file=sys.stderr)


def spmm_raw(m, n, nnz, rp, ci, a, alpha, X, ldX, beta, Y, ldY, nrhs, NRHS=0, job_ptr=None, blocks=0,
             stream=None):
    """Direct tg_csr_spmm on torch CUDA tensors; returns the status (no raise for the
    reference's no-op cases TG_ERR_ALIAS / TG_ERR_NRHS)."""
    rp_type = capi.I64 if rp.dtype == torch.int64 else capi.I32
    val_type = _TORCH_TO_TG[Y.dtype]
    st = capi.lib().tg_csr_spmm(
        m, n, nnz, rp.data_ptr(), rp_type, ci.data_ptr() if ci is not None and ci.numel() else None,
        a.data_ptr() if a is not None else None, val_type, float(alpha),
        X.data_ptr() if X is not None else None, ldX, float(beta), Y.data_ptr(), ldY, nrhs, NRHS,
        job_ptr.data_ptr() if job_ptr is not None else None, blocks, stream_handle(stream))
    if st in (capi.ERR_ALIAS, capi.ERR_NRHS):
        return st
    return capi.check(st, "tg_csr_spmm")


class SlicePlan:
    """tg_csr_plan_t of one device CSR pattern (column-sliced SpMM for small X)."""

    def __init__(self, m, n, nnz, rp, ci, stream=None):
        self.h = C.c_void_p()
        self._keep = (rp, ci)
        capi.call("tg_csr_plan_create", m, n, nnz, rp.data_ptr(), capi.I64 if rp.dtype == torch.int64 else capi.I32,
                  ci.data_ptr() if ci.numel() else None, C.byref(self.h), stream_handle(stream))

    def slices(self, dtype, nrhs, a=None):
        """Slices a call with values `a` uses (8 on the re-laid layout)."""
        if a is not None and self.uses_split(a, nrhs):
            return 8
        return capi.lib().tg_csr_plan_slices(self.h, _TORCH_TO_TG[dtype], nrhs)

    def set_values(self, a, stream=None):
        """tg_csr_plan_set_values: keep a re-laid copy of the values `a` (a CUDA tensor that
        must not change afterwards without another set_values); None drops it for a's type."""
        self._vals = getattr(self, "_vals", {})
        self._vals[a.dtype] = a
        capi.call("tg_csr_plan_set_values", self.h, a.data_ptr(), _TORCH_TO_TG[a.dtype], stream_handle(stream))

    def ensure_split(self, a, nrhs, n):
        """Registers the matrix's own values `a` on the first call whose shape takes the
        re-laid layout (nrhs <= 4, X of 2-64 MB), as the drop-in MatrixCSR does."""
        xbytes = n * nrhs * a.element_size()
        vals = getattr(self, "_vals", {})
        if nrhs <= 4 and (2 << 20) < xbytes <= (64 << 20) and vals.get(a.dtype) is not a:
            self.set_values(a)

    def long_rows(self):
        """Rows split off the row kernels (more than the spmm_long_row knob's nonzeros)."""
        return int(capi.lib().tg_csr_plan_long_rows(self.h))

    def uses_split(self, a, nrhs):
        return bool(capi.lib().tg_csr_plan_uses_split(self.h, a.data_ptr(), _TORCH_TO_TG[a.dtype], nrhs))

    def spmm(self, a, alpha, X, ldX, beta, Y, ldY, nrhs, NRHS=0, stream=None):
        """tg_csr_spmm_plan; returns the status like spmm_raw."""
        st = capi.lib().tg_csr_spmm_plan(
            self.h, a.data_ptr() if a is not None else None, _TORCH_TO_TG[Y.dtype], float(alpha),
            X.data_ptr() if X is not None else None, ldX, float(beta), Y.data_ptr(), ldY, nrhs, NRHS,
            stream_handle(stream))
        if st in (capi.ERR_ALIAS, capi.ERR_NRHS):
            return st
        return capi.check(st, "tg_csr_spmm_plan")

    def __del__(self):
        try:
            if self.h:
                capi.lib().tg_csr_plan_destroy(self.h)
        except Exception:
            pass


class MatrixCSR:
    """m x n CSR matrix: ``outer`` (m+1, int32/int64), ``inner`` (nnz, int32), ``values``
    (nnz, float64/float32, or None for a pattern/binary matrix).

    Mirrors MatrixCSR<Scal,Int,LInt> (src/Sparse/MatrixCSR.hpp:9-10, 87-100) and its Dot
    overloads (src/Sparse/MatrixCSR.hpp:965-1086)."""

    def __init__(self, outer, inner, values, m, n, thread_count=1, device="cuda"):
        self.m, self.n = int(m), int(n)
        self.outer = np.ascontiguousarray(outer)
        self.inner = np.ascontiguousarray(inner, dtype=np.int32)
        self.values = None if values is None else np.ascontiguousarray(values)
        self.thread_count = int(thread_count)
        self.device = device
        self._dev = None
        self._job_ptr = None
        self._plan = None

    @classmethod
    def from_device(cls, rp, ci, vals, m, n):
        """Wraps CUDA tensors that already hold the CSR arrays (host copies are made for
        the accessors; the device arrays are used as they are)."""
        A = cls(rp.cpu().numpy(), ci.cpu().numpy(), None if vals is None else vals.cpu().numpy(), m, n,
                device=rp.device)
        A._dev = {"rp": rp, "ci": ci, "vals": {} if vals is None else {vals.dtype: vals}}
        return A

    # -- MatrixMarket files (MatrixCSR/MatrixMarket.hpp:3-306) ----------------------
    @classmethod
    def from_matrix_market(cls, path, thread_count=1, dtype=np.float64, threads=0):
        """LoadFromMatrixMarket: the file's entries are parsed on the host (tg_mm_read) and
        assembled on the GPU (tg_csr_from_triples: duplicates summed, off-diagonal entries
        mirrored for "symmetric" files).  Values of "pattern" files are 1."""
        from . import structure
        m, n, cnt, field, symm = C.c_int32(0), C.c_int32(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
        p = os.fsencode(path)
        capi.call("tg_mm_info", p, C.byref(m), C.byref(n), C.byref(cnt), C.byref(field), C.byref(symm))
        k = cnt.value
        I = np.empty(max(k, 1), np.int32)
        J = np.empty(max(k, 1), np.int32)
        V = np.empty(max(k, 1), np.float64)
        capi.call("tg_mm_read", p, k, I.ctypes.data, J.ctypes.data, V.ctypes.data, int(threads))
        rp, ci, val = structure.from_triples(I[:k], J[:k], V[:k], m.value, n.value, compress=True,
                                             symmetrize=bool(symm.value))
        return cls(rp, ci, val.astype(dtype), m.value, n.value, thread_count)

    def write_matrix_market(self, path, threads=0):
        """WriteToMatrixMarket: "coordinate real general" ("pattern" without values), one
        line per nonzero in row order."""
        rp = np.ascontiguousarray(self.outer)
        vals = None if self.values is None else np.ascontiguousarray(self.values)
        capi.call("tg_mm_write", os.fsencode(path), self.m, self.n, rp.ctypes.data,
                  capi.I64 if rp.dtype == np.int64 else capi.I32, self.inner.ctypes.data if self.inner.size else None,
                  None if vals is None else vals.ctypes.data,
                  capi.F32 if vals is not None and vals.dtype == np.float32 else capi.F64, int(threads))

    # -- accessors (PatternCSR.hpp / MatrixCSR.hpp) --------------------------------
    def RowCount(self): return self.m
    def ColCount(self): return self.n
    def NonzeroCount(self): return int(self.outer[self.m]) if self.m >= 0 else 0
    def ThreadCount(self): return self.thread_count
    def Outer(self): return self.outer
    def Inner(self): return self.inner
    def Values(self): return self.values

    def SetThreadCount(self, t):
        # the GPU path ignores the host thread count; kept for API parity
        self.thread_count = int(t)

    def WellFormedQ(self):  # PatternCSR.hpp:1146-1167
        return self.outer.size > 1 and int(self.outer[-1]) > 0

    # -- device mirror -------------------------------------------------------------
    def device_arrays(self, dtype=None):
        """(rp, ci, values) torch CUDA tensors; values converted to `dtype` if given."""
        if self._dev is None:
            rp = torch.from_numpy(self.outer).to(self.device)
            ci = torch.from_numpy(self.inner).to(self.device)
            vals = None if self.values is None else torch.from_numpy(self.values).to(self.device)
            self._dev = {"rp": rp, "ci": ci, "vals": {}}
            if vals is not None:
                self._dev["vals"][vals.dtype] = vals
        vals = None
        if self.values is not None:
            cache = self._dev["vals"]
            want = dtype if dtype is not None else next(iter(cache))
            if want not in cache:
                cache[want] = next(iter(cache.values())).to(want)
            vals = cache[want]
        return self._dev["rp"], self._dev["ci"], vals

    def Plan(self):
        """Column-slice plan of the device pattern (tg_csr_plan_create), made on first use."""
        if self._plan is None:
            rp, ci, _ = self.device_arrays()
            self._plan = SlicePlan(self.m, self.n, self.NonzeroCount(), rp, ci)
        return self._plan

    def JobPtr(self, nrhs=1):
        """Device row partition (tg_csr_job_pointers), cached per nrhs."""
        if self._job_ptr is None:
            self._job_ptr = {}
        if nrhs not in self._job_ptr:
            rp, _, _ = self.device_arrays()
            blocks = capi.lib().tg_csr_default_blocks(self.m, self.NonzeroCount(), nrhs)
            jp = torch.empty(max(blocks, 0) + 1, dtype=torch.int32, device=self.device)
            if blocks > 0:
                capi.call("tg_csr_job_pointers", self.m, rp.data_ptr(),
                          capi.I64 if rp.dtype == torch.int64 else capi.I32, blocks, jp.data_ptr(),
                          stream_handle())
            self._job_ptr[nrhs] = (jp, blocks)
        return self._job_ptr[nrhs]

    # -- products ----------------------------------------------------------------
    def Dot(self, alpha, X, beta, Y, nrhs=None, NRHS=0, ext_values=None, use_job_ptr=False, stream=None):
        """Y <- alpha * A X + beta * Y on CUDA tensors.  1-D X/Y: the Tensor1 overload
        (MatrixCSR.hpp:987-1003); 2-D: the Tensor2 overload (:1006-1022) with nrhs =
        X.shape[1].  A shape mismatch prints the reference's message and does nothing."""
        if X.dim() == 1:
            if X.shape[0] != self.n or Y.shape[0] != self.m:
                _eprint(self._name("Dot") + ": shapes of matrix, input, and output do not match.")
                return capi.ERR_SHAPE
            k, ldX, ldY = 1, 1, 1
        else:
            if X.shape[0] != self.n or Y.shape[0] != self.m or X.shape[1] != Y.shape[1]:
                _eprint(self._name("Dot") + ": shapes of matrix, input, and output do not match.")
                return capi.ERR_SHAPE
            k = X.shape[1] if nrhs is None else int(nrhs)
            ldX, ldY = X.stride(0), Y.stride(0)
        if not self.WellFormedQ():
            _eprint(self._name("Dot_") + ": Not WellFormedQ(). Doing nothing.")
        rp, ci, vals = self.device_arrays(Y.dtype)
        if ext_values is not None:
            vals = ext_values
        if use_job_ptr:
            jp, blocks = self.JobPtr(k)
            st = spmm_raw(self.m, self.n, self.NonzeroCount(), rp, ci, vals, alpha, X, ldX, beta, Y, ldY, k,
                          NRHS, jp, blocks, stream)
        else:
            plan = self.Plan()
            if ext_values is None and vals is not None:
                plan.ensure_split(vals, k, self.n)
            st = plan.spmm(vals, alpha, X, ldX, beta, Y, ldY, k, NRHS, stream)
        if st == capi.ERR_ALIAS:
            _eprint(self._name("Dot") + ": Input and output pointer coincide. This is not safe. Aborting.")
        elif st == capi.ERR_NRHS:
            _eprint(self._name("Dot") + ": nrhs != NRHS. Doing nothing.")
        return st

    def _name(self, method):
        vt = "Real64" if self.values is None or self.values.dtype == np.float64 else "Real32"
        it = "Int64" if self.outer.dtype == np.int64 else "Int32"
        return f"Sparse::MatrixCSR<{vt},Int32,{it}>::{method}"


class KernelMatrixCSR:
    """Sparse::KernelMatrixCSR<Kernel_T> (KernelMatrixCSR.hpp:7-227) with the dense block kernel
    DenseBlockKernel_Tiny<ROWS, COLS, NRHS, ..., a_RM, ..., x_RM, ..., y_RM>
    (BlockKernels/DenseBlockKernel_Tiny.hpp): a pattern of mb x nb blocks (outer, inner) whose
    nonzeros are ROWS x COLS dense blocks.  The block values are passed to Dot, as in the
    reference (Dot(A, alpha, X, beta, Y, nrhs)); X holds nb chunks of COLS x nrhs, Y mb chunks
    of ROWS x nrhs, each row-major (x_RM / y_RM) or transposed.  The product runs on the GPU
    (tg_bsr_spmm)."""

    def __init__(self, outer, inner, mb, nb, rows, cols, a_row_major=True, x_row_major=True, y_row_major=True,
                 thread_count=1, device="cuda"):
        self.mb, self.nb = int(mb), int(nb)
        self.ROWS, self.COLS = int(rows), int(cols)
        self.a_rm, self.x_rm, self.y_rm = bool(a_row_major), bool(x_row_major), bool(y_row_major)
        self.outer = np.ascontiguousarray(outer)
        self.inner = np.ascontiguousarray(inner, dtype=np.int32)
        self.thread_count = int(thread_count)
        self.rp = torch.from_numpy(self.outer).to(device)
        self.ci = torch.from_numpy(self.inner).to(device)

    def RowCount(self): return self.mb * self.ROWS
    def ColCount(self): return self.nb * self.COLS
    def BlockNonzeroCount(self): return int(self.outer[self.mb])
    def NonzeroCount(self): return self.BlockNonzeroCount() * self.ROWS * self.COLS

    def Dot(self, A, alpha, X, beta, Y, nrhs, stream=None):
        """Y <- alpha * A X + beta * Y for block values A (CUDA tensor, nnzb*ROWS*COLS) and CUDA
        X (nb*COLS*nrhs entries), Y (mb*ROWS*nrhs entries) of A's dtype."""
        if X.numel() != self.nb * self.COLS * nrhs or Y.numel() != self.mb * self.ROWS * nrhs:
            _eprint(self._name("Dot") + ": shapes of matrix, input, and output do not match.")
            return capi.ERR_SHAPE
        st = capi.lib().tg_bsr_spmm(
            self.mb, self.nb, self.BlockNonzeroCount(), self.ROWS, self.COLS, self.rp.data_ptr(),
            capi.I64 if self.rp.dtype == torch.int64 else capi.I32,
            self.ci.data_ptr() if self.ci.numel() else None, A.data_ptr() if A.numel() else None,
            _TORCH_TO_TG[Y.dtype], int(self.a_rm), float(alpha), X.data_ptr(), int(self.x_rm), float(beta),
            Y.data_ptr(), int(self.y_rm), int(nrhs), stream_handle(stream))
        if st == capi.ERR_ALIAS:
            _eprint(self._name("Dot") + ": Input and output pointer coincide. Doing nothing.")
            return st
        return capi.check(st, "tg_bsr_spmm")

    def FillLowerTriangleFromUpperTriangle(self, A, stream=None):
        """Every block (i, j), j < i, becomes the transpose of block (j, i) (square blocks,
        sorted block columns; KernelMatrixCSR.hpp:71-128)."""
        if self.ROWS != self.COLS:
            raise ValueError("FillLowerTriangleFromUpperTriangle needs square blocks")
        capi.call("tg_bsr_fill_lower", self.mb, self.rp.data_ptr(),
                  capi.I64 if self.rp.dtype == torch.int64 else capi.I32,
                  self.ci.data_ptr() if self.ci.numel() else None, self.ROWS, A.data_ptr(), _TORCH_TO_TG[A.dtype],
                  stream_handle(stream))

    def _name(self, method):
        return f"Sparse::KernelMatrixCSR<DenseBlockKernel_Tiny<{self.ROWS},{self.COLS},...>>::{method}"
