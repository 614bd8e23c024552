"""tg_cg_solve_noalloc (the CG solve whose workspace is the library's static device
array) against tg_cg_solve with an allocated workspace: the same kernels on the same
inputs, so the results must agree bit for bit -- on two calls in a row, the second with a
smaller system (the array keeps the first call's contents), in FP64 and FP32, with and
without the Jacobi diagonal.  A system too large for the array is refused before any
device work.  Everything goes through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from tensors_amd import _capi as capi
from tensors_amd import generators as gen

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20  # TG_CG_NOALLOC_SCRATCH_BYTES


def system(g, k, dtype, prec, dev, seed):
    rp, ci, val = oracle.grid_laplacian(g, 0.1)
    n = g * g
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    B = gen.uniform(seed, n * k).reshape(n, k)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    return (n, len(ci), T(rp, torch.int32), T(ci, torch.int32), T(val, dtype), None if d is None else T(d, dtype),
            T(B, dtype))


def solve(entry, sysd, k, dtype, tol, max_iter, dev):
    n, nnz, rp, ci, val, d, B = sysd
    tg = capi.F64 if dtype == torch.float64 else capi.F32
    X = torch.full((n, k), 7.0, dtype=dtype, device=dev)  # overwritten: b = 0, no initial guess
    ok, it = C.c_int32(0), C.c_int32(0)
    rr = np.zeros(k)
    head = (n, nnz, rp.data_ptr(), capi.I32, ci.data_ptr(), val.data_ptr(), tg, 0 if d is None else 1,
            None if d is None else d.data_ptr(), k, max_iter, 1.0, B.data_ptr(), k, 0.0, X.data_ptr(), k, tol, 0)
    if entry == "tg_cg_solve":
        prec_kind = 0 if d is None else 1
        nb = capi.lib().tg_cg_workspace_bytes(n, k, tg, prec_kind)
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        capi.call(entry, *head, ws.data_ptr(), nb, C.byref(ok), C.byref(it), rr.ctypes.data, None)
    else:
        capi.call(entry, *head, C.byref(ok), C.byref(it), rr.ctypes.data, None)
    torch.cuda.synchronize()
    return bool(ok.value), it.value, rr, X.cpu()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("prec", [False, True], ids=["identity", "jacobi"])
def test_noalloc_matches_allocated_entry(cuda, dtype, prec):
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    # first call: 96^2 grid, 8 right-hand sides; second call: a smaller 20^2 grid, 3 right-hand sides
    for g, k, seed in ((96, 8, 61), (20, 3, 62)):
        sysd = system(g, k, dtype, prec, cuda, seed)
        print(f"noalloc vs tg_cg_solve: grid {g} nrhs {k} {dtype} prec {prec}", flush=True)
        ok_a, it_a, rr_a, X_a = solve("tg_cg_solve", sysd, k, dtype, tol, 2000, cuda)
        ok_s, it_s, rr_s, X_s = solve("tg_cg_solve_noalloc", sysd, k, dtype, tol, 2000, cuda)
        assert ok_a and ok_s and it_a == it_s > 0
        assert torch.equal(X_a, X_s)
        assert np.array_equal(rr_a, rr_s)


def test_noalloc_refuses_oversize_system(cuda):
    n, k = 1 << 20, 8  # 8 MiB per FP64 vector column block x 8 columns: far above the array
    tg = capi.F64
    assert capi.lib().tg_cg_workspace_bytes(n, k, tg, 0) > SCRATCH
    dummy = torch.zeros(16, dtype=torch.float64, device=cuda)
    ok, it = C.c_int32(5), C.c_int32(5)
    st = capi.lib().tg_cg_solve_noalloc(n, 1, dummy.data_ptr(), capi.I32, dummy.data_ptr(), dummy.data_ptr(), tg, 0,
                                         None, k, 10, 1.0, dummy.data_ptr(), k, 0.0, dummy.data_ptr(), k, 1e-8, 0,
                                         C.byref(ok), C.byref(it), None, None)
    assert st == capi.ERR_WORKSPACE and "static scratch" in capi.last_error()
    assert ok.value == 0 and it.value == 0
