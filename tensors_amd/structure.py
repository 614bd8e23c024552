"""Structure work on the GPU via the C ABI (tg_csr_from_triples / tg_csr_transpose /
tg_csr_spgemm): host numpy in, host numpy out (int64 row pointers)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi

_DEV = "cuda"


def _t(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(_DEV)
    return t if dtype is None else t.to(dtype)


def from_triples(I, J, V, m, n, compress=True, symmetrize=False):
    """Triples -> CSR on the GPU (MatrixCSR::FromTriples semantics).  V may be None."""
    count = len(I)
    cap = max(1, count * (2 if symmetrize else 1))
    dI, dJ = _t(np.asarray(I, np.int32)), _t(np.asarray(J, np.int32))
    dV = None if V is None else _t(np.asarray(V, np.float64))
    rp = torch.empty(m + 1, dtype=torch.int64, device=_DEV)
    ci = torch.empty(cap, dtype=torch.int32, device=_DEV)
    va = torch.empty(cap, dtype=torch.float64, device=_DEV)
    nb = capi.lib().tg_csr_from_triples_workspace_bytes(count, m, int(symmetrize))
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=_DEV)
    nnz = C.c_int64(0)
    capi.call("tg_csr_from_triples", count, dI.data_ptr() if count else None, dJ.data_ptr() if count else None,
              dV.data_ptr() if dV is not None and count else None, capi.F64, m, n, int(compress), int(symmetrize),
              rp.data_ptr(), capi.I64, ci.data_ptr(), va.data_ptr(), cap, ws.data_ptr(), nb, C.byref(nnz), None)
    k = nnz.value
    return rp.cpu().numpy(), ci[:k].cpu().numpy(), (va[:k].cpu().numpy() if V is not None else None)


def transpose(m, n, rp, ci, val):
    nnz = int(rp[m])
    drp, dci = _t(np.asarray(rp, np.int64)), _t(np.asarray(ci, np.int32))
    dval = None if val is None else _t(np.asarray(val, np.float64))
    orp = torch.empty(n + 1, dtype=torch.int64, device=_DEV)
    oci = torch.empty(max(nnz, 1), dtype=torch.int32, device=_DEV)
    ova = torch.empty(max(nnz, 1), dtype=torch.float64, device=_DEV)
    nb = capi.lib().tg_csr_transpose_workspace_bytes(m, n, nnz)
    ws = torch.empty(nb, dtype=torch.uint8, device=_DEV)
    capi.call("tg_csr_transpose", m, n, nnz, drp.data_ptr(), capi.I64, dci.data_ptr(),
              dval.data_ptr() if dval is not None else None, capi.F64, orp.data_ptr(), oci.data_ptr(),
              ova.data_ptr() if dval is not None else None, ws.data_ptr(), nb, None)
    return orp.cpu().numpy(), oci[:nnz].cpu().numpy(), (ova[:nnz].cpu().numpy() if val is not None else None)


def spgemm(m, k, n, A, B, values=True):
    Arp, Aci, Aval = A
    Brp, Bci, Bval = B
    dArp, dAci = _t(np.asarray(Arp, np.int64)), _t(np.asarray(Aci, np.int32))
    dBrp, dBci = _t(np.asarray(Brp, np.int64)), _t(np.asarray(Bci, np.int32))
    dAv = _t(np.asarray(Aval, np.float64)) if values else None
    dBv = _t(np.asarray(Bval, np.float64)) if values else None
    w0 = torch.empty(16 * (m + 1) + (1 << 20), dtype=torch.uint8, device=_DEV)
    E = C.c_int64(0)
    capi.call("tg_csr_spgemm_expanded_count", m, dArp.data_ptr(), dAci.data_ptr(), dBrp.data_ptr(), capi.I64,
              C.byref(E), w0.data_ptr(), w0.numel(), None)
    e = E.value
    nb = capi.lib().tg_csr_spgemm_workspace_bytes(m, n, e, capi.F64)
    ws = torch.empty(nb, dtype=torch.uint8, device=_DEV)
    Crp = torch.empty(m + 1, dtype=torch.int64, device=_DEV)
    Cci = torch.empty(max(e, 1), dtype=torch.int32, device=_DEV)
    Cva = torch.empty(max(e, 1), dtype=torch.float64, device=_DEV)
    nnz = C.c_int64(0)
    capi.call("tg_csr_spgemm", m, k, n, dArp.data_ptr(), dAci.data_ptr(), dAv.data_ptr() if values else None,
              dBrp.data_ptr(), dBci.data_ptr(), dBv.data_ptr() if values else None, capi.I64, capi.F64, e,
              Crp.data_ptr(), Cci.data_ptr(), Cva.data_ptr(), ws.data_ptr(), nb, C.byref(nnz), None)
    z = nnz.value
    return Crp.cpu().numpy(), Cci[:z].cpu().numpy(), (Cva[:z].cpu().numpy() if values else None)
