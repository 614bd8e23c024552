"""Python handle over the batched CG engine (tg_batch_*): many SPD systems of uneven size
advanced together, continuously fed."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi
from .sparse import stream_handle


class BatchCG:
    """Holds a tg_batch_t.  add() takes device tensors that must stay alive until the
    system finished (the object keeps references until release)."""

    def __init__(self, stream=None):
        self.h = C.c_void_p()
        self.stream = stream
        capi.call("tg_batch_create", C.byref(self.h), stream_handle(stream))
        self._keep = {}

    def close(self):
        if self.h:
            capi.lib().tg_batch_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, rp, ci, val, B, X, tol, max_iter, diag=None, mixed=False, true_residual=False):
        """rp/ci int32 CUDA tensors, val/diag/B/X float64 CUDA tensors (X written in place).
        mixed: FP32 iteration with FP64 residual replacement, stopping on the true residual;
        true_residual (FP64 only): the same replacement and stop in FP64 (mode 2) instead of
        the reference's preconditioned test."""
        n = rp.numel() - 1
        k = B.shape[1] if B.dim() == 2 else 1
        sid = C.c_int64(0)
        mode = 1 if mixed else (2 if true_residual else 0)
        capi.call("tg_batch_add_cg", self.h, n, ci.numel(), rp.data_ptr(), ci.data_ptr(), val.data_ptr(),
                  diag.data_ptr() if diag is not None else None, k, B.data_ptr(), X.data_ptr(), float(tol),
                  int(max_iter), mode, C.byref(sid))
        self._keep[sid.value] = (rp, ci, val, diag, B, X, k)
        return sid.value

    def step(self, iterations=8):
        capi.call("tg_batch_step", self.h, int(iterations))

    def finished(self, cap=4096):
        ids = np.zeros(cap, np.int64)
        cnt = C.c_int32(0)
        capi.call("tg_batch_finished", self.h, ids.ctypes.data, cap, C.byref(cnt))
        return [int(v) for v in ids[: cnt.value]]

    def result(self, sid):
        k = self._keep[sid][6]
        fin, ok, it, outer = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        rr = np.zeros(k)
        capi.call("tg_batch_result", self.h, sid, C.byref(fin), C.byref(ok), C.byref(it), C.byref(outer),
                  rr.ctypes.data)
        return {"finished": bool(fin.value), "succeeded": bool(ok.value), "iterations": it.value,
                "outer": outer.value, "rel_res": rr}

    def release(self, sid):
        capi.call("tg_batch_release", self.h, sid)
        self._keep.pop(sid, None)

    def active(self):
        n, b = C.c_int32(0), C.c_int64(0)
        capi.call("tg_batch_active", self.h, C.byref(n), C.byref(b))
        return n.value, b.value

    def run(self, iterations=8, max_steps=1_000_000):
        """Steps until every added system finished; returns the finished ids in order."""
        done = []
        for _ in range(max_steps):
            active, _ = self.active()
            if active == 0:
                break
            self.step(iterations)
            done += self.finished()
        else:
            raise capi.TensorsError(capi.ERR_INVALID, "BatchCG.run", "systems still active after max_steps steps")
        done += self.finished()
        return done


class BatchGMRES:
    """Holds a tg_gbatch_t: GMRES(restart) systems advanced together in lock-step
    restart cycles.  add() takes device tensors that must stay alive until the system
    finished (the object keeps references until release)."""

    def __init__(self, stream=None):
        self.h = C.c_void_p()
        self.stream = stream
        capi.call("tg_gbatch_create", C.byref(self.h), stream_handle(stream))
        self._keep = {}

    def close(self):
        if self.h:
            capi.lib().tg_gbatch_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, rp, ci, val, B, X, tol, restart=30, max_restarts=100, diag=None):
        """rp/ci int32 CUDA tensors, val/diag/B/X float64 CUDA tensors (X written in place);
        diag: left diagonal preconditioner (z = diag * r) or None."""
        n = rp.numel() - 1
        k = B.shape[1] if B.dim() == 2 else 1
        sid = C.c_int64(0)
        capi.call("tg_gbatch_add", self.h, n, ci.numel(), rp.data_ptr(), ci.data_ptr(), val.data_ptr(),
                  diag.data_ptr() if diag is not None else None, k, B.data_ptr(), X.data_ptr(), float(tol),
                  int(restart), int(max_restarts), C.byref(sid))
        self._keep[sid.value] = (rp, ci, val, diag, B, X, k)
        return sid.value

    def step(self, cycles=1):
        capi.call("tg_gbatch_step", self.h, int(cycles))

    def finished(self, cap=4096):
        ids = np.zeros(cap, np.int64)
        cnt = C.c_int32(0)
        capi.call("tg_gbatch_finished", self.h, ids.ctypes.data, cap, C.byref(cnt))
        return [int(v) for v in ids[: cnt.value]]

    def result(self, sid):
        k = self._keep[sid][6]
        fin, ok, it, rs = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        rr = np.zeros(k)
        capi.call("tg_gbatch_result", self.h, sid, C.byref(fin), C.byref(ok), C.byref(it), C.byref(rs),
                  rr.ctypes.data)
        return {"finished": bool(fin.value), "succeeded": bool(ok.value), "iterations": it.value,
                "restarts": rs.value, "rel_res": rr}

    def release(self, sid):
        capi.call("tg_gbatch_release", self.h, sid)
        self._keep.pop(sid, None)

    def active(self):
        n, b = C.c_int32(0), C.c_int64(0)
        capi.call("tg_gbatch_active", self.h, C.byref(n), C.byref(b))
        return n.value, b.value

    def run(self, cycles=1, max_steps=100000):
        """Steps until every added system finished; returns the finished ids in order."""
        done = []
        for _ in range(max_steps):
            active, _ = self.active()
            if active == 0:
                break
            self.step(cycles)
            done += self.finished()
        else:
            raise capi.TensorsError(capi.ERR_INVALID, "BatchGMRES.run", "systems still active after max_steps steps")
        done += self.finished()
        return done
