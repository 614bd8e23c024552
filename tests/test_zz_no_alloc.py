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


def lap3d_system(nx, ny, nz, k, dtype, dev, seed):
    """7-point Laplacian (+ 0.1) of an nx x ny x nz grid, x fastest: band nx * ny."""
    def path(m):
        return sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])

    A = (sp.kron(sp.identity(nz * ny), path(nx)) + sp.kron(sp.kron(sp.identity(nz), path(ny)), sp.identity(nx))
         + sp.kron(path(nz), sp.identity(ny * nx)) + 0.1 * sp.identity(nx * ny * nz)).tocsr()
    A.sort_indices()
    n = A.shape[0]
    B = gen.uniform(seed, n * k).reshape(n, k)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    return (n, A.nnz, T(A.indptr, torch.int32), T(A.indices, torch.int32), T(A.data, dtype), None, T(B, dtype))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_noalloc_matches_allocated_entry_in_panel_order(cuda, dtype):
    """A system large enough for the band-dependent SpMM row order (65 536 rows, band 8192:
    panels of one plane): both entries measure the band (the no-allocation one into the
    array's last 4 KiB) and take the same order, so they still agree bit for bit; then a
    smaller system.  In FP64 the first system's result does depend on that order (knob
    cg_panel), which shows the case takes the band-dependent path; in FP32 the products
    of two floats are exact in the double partial sums, so the order need not show."""
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    for sysd, k, what in ((lap3d_system(128, 64, 8, 4, dtype, cuda, 71), 4, "128x64x8"),
                          (system(40, 2, dtype, False, cuda, 72), 2, "grid 40")):
        print(f"noalloc vs tg_cg_solve: {what} nrhs {k} {dtype}", flush=True)
        ok_a, it_a, rr_a, X_a = solve("tg_cg_solve", sysd, k, dtype, tol, 3000, cuda)
        ok_s, it_s, rr_s, X_s = solve("tg_cg_solve_noalloc", sysd, k, dtype, tol, 3000, cuda)
        assert ok_a and ok_s and it_a == it_s > 0
        assert torch.equal(X_a, X_s)
        assert np.array_equal(rr_a, rr_s)
        if what == "128x64x8" and dtype == torch.float64:
            before = int(capi.lib().tg_tune_get(b"cg_panel"))
            capi.call("tg_tune_set", b"cg_panel", 0)
            try:
                _, _, _, X_c = solve("tg_cg_solve_noalloc", sysd, k, dtype, tol, 3000, cuda)
            finally:
                capi.call("tg_tune_set", b"cg_panel", before)
            assert not torch.equal(X_c, X_s)


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
