"""Python handle over the serving loop (tg_server_*): CG and GMRES requests with host
arrays, admitted under a device-memory budget and solved with their work overlapping."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Server:
    """tg_server_t.  Requests keep references to their host arrays until release()."""

    def __init__(self, budget_bytes=200 << 30, gmres_workers=3, cg_iterations_per_poll=16):
        self.h = C.c_void_p()
        capi.call("tg_server_create", int(budget_bytes), int(gmres_workers), int(cg_iterations_per_poll),
                  C.byref(self.h))
        self._req = {}

    def close(self):
        if self.h:
            capi.lib().tg_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def request_bytes(kind, n, nnz, nrhs, has_diag=False, mixed=False, restart=30):
        return capi.lib().tg_server_request_bytes(0 if kind == "cg" else 1, n, nnz, nrhs, int(has_diag),
                                                  int(mixed), restart)

    def submit_cg(self, rp, ci, val, B, tol=1e-10, max_iter=20000, inv_diag=None, mixed=True):
        rp, ci, val = _c(rp, np.int32), _c(ci, np.int32), _c(val, np.float64)
        B = _c(B, np.float64)
        n = len(rp) - 1
        k = 1 if B.ndim == 1 else B.shape[1]
        X = np.zeros((n, k))
        d = None if inv_diag is None else _c(inv_diag, np.float64)
        sid = C.c_int64(0)
        capi.call("tg_server_submit_cg", self.h, n, len(ci), rp.ctypes.data, ci.ctypes.data, val.ctypes.data,
                  d.ctypes.data if d is not None else None, k, B.ctypes.data, X.ctypes.data, float(tol),
                  int(max_iter), int(bool(mixed)), C.byref(sid))
        self._req[sid.value] = (rp, ci, val, d, B, X, k)
        return sid.value

    def submit_gmres(self, rp, ci, val, B, tol=1e-8, restart=30, max_restarts=100):
        rp, ci, val = _c(rp, np.int32), _c(ci, np.int32), _c(val, np.float64)
        B = _c(B, np.float64)
        n = len(rp) - 1
        k = 1 if B.ndim == 1 else B.shape[1]
        X = np.zeros((n, k))
        sid = C.c_int64(0)
        capi.call("tg_server_submit_gmres", self.h, n, len(ci), rp.ctypes.data, ci.ctypes.data, val.ctypes.data,
                  k, B.ctypes.data, X.ctypes.data, float(tol), int(restart), int(max_restarts), C.byref(sid))
        self._req[sid.value] = (rp, ci, val, None, B, X, k)
        return sid.value

    def wait(self, timeout_s=0.05, cap=4096):
        ids = np.zeros(cap, np.int64)
        cnt = C.c_int32(0)
        capi.call("tg_server_wait", self.h, ids.ctypes.data, cap, C.byref(cnt), float(timeout_s))
        return [int(v) for v in ids[: cnt.value]]

    def result(self, sid):
        k = self._req[sid][6]
        st, ok, it, rs = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        rr = np.zeros(k)
        tm = np.zeros(3)
        capi.call("tg_server_result", self.h, sid, C.byref(st), C.byref(ok), C.byref(it), C.byref(rs),
                  rr.ctypes.data, tm.ctypes.data)
        return {"status": st.value, "succeeded": bool(ok.value), "iterations": it.value, "restarts": rs.value,
                "rel_res": rr, "t_submit": tm[0], "t_admit": tm[1], "t_done": tm[2], "X": self._req[sid][5]}

    def release(self, sid):
        capi.call("tg_server_release", self.h, sid)
        self._req.pop(sid, None)

    def stats(self):
        a, b, q, r = C.c_int64(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
        capi.call("tg_server_stats", self.h, C.byref(a), C.byref(b), C.byref(q), C.byref(r))
        return {"bytes_in_use": a.value, "peak_bytes": b.value, "queued": q.value, "running": r.value}

    def drain(self, ids, timeout_s=600.0):
        """Waits until every id in `ids` finished; returns {id: result}."""
        import time

        want, out = set(ids), {}
        t_end = time.monotonic() + timeout_s
        while want - set(out):
            if time.monotonic() > t_end:
                raise capi.TensorsError("Server.drain: timed out")
            for sid in self.wait(0.05):
                if sid in want:
                    out[sid] = self.result(sid)
        return out
