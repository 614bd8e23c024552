"""GPU parity of tg_cg_solve (block PCG behind ConjugateGradient::operator()) against the
oracle's restatement of ConjugateGradient.hpp and against dense solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import _capi as capi
from tensors_amd import generators as gen
from tensors_amd.solvers import ConjugateGradient, cg_solve_mixed
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu


def dev(a, d, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dtype)


def grid_system(g, mass=1.0):
    rp, ci, val = oracle.grid_laplacian(g, mass)
    return rp, ci, val, g * g


@pytest.fixture
def fused_knob():
    """Restores the process-wide cg_fused knob after the test."""
    before = int(capi.lib().tg_tune_get(b"cg_fused"))
    yield lambda v: capi.call("tg_tune_set", b"cg_fused", int(v))
    capi.call("tg_tune_set", b"cg_fused", before)


@pytest.mark.parametrize("k", [1, 2, 5, 16, 32])
@pytest.mark.parametrize("prec", [False, True])
def test_cg_matches_oracle(cuda, k, prec):
    rp, ci, val, n = grid_system(24, 0.1)
    B = gen.uniform(3003 + k, n * k).reshape(n, k)
    d = None
    if prec:
        D = sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
        d = 1.0 / D
    Xo = np.zeros((n, k))
    ok_o, it_o, rr_o = oracle.cg(n, rp, ci, val, B, Xo, k, 1000, 1e-10, prec_diag=d)
    A = MatrixCSR(rp, ci, val, n, n)
    cg = ConjugateGradient(n, 1000, k)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok = cg(A, None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, 1e-10)
    assert ok and ok_o == 1
    assert abs(cg.IterationCount() - it_o) <= 1
    X = Xd.cpu().numpy()
    assert rel_err(X, Xo) < 1e-8
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    res = np.abs(B - Dm @ X).max() / np.abs(B).max()
    assert res < 1e-8
    assert np.all(cg.RelativeResiduals() <= 1e-10)


@pytest.mark.parametrize("a,b", [(-1.0, 0.0), (1.0, 1.0), (np.pi, -1.0), (0.0, np.pi)])
def test_cg_alpha_beta(cuda, a, b):
    rp, ci, val, n = grid_system(12, 0.5)
    k = 3
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (n, k))
    X0 = rng.uniform(-1, 1, (n, k))
    Xo = X0.copy()
    oracle.cg(n, rp, ci, val, B, Xo, k, 500, 1e-12, a=a, b=b)
    cg = ConjugateGradient(n, 500, k)
    Xd = dev(X0, cuda)
    assert cg(MatrixCSR(rp, ci, val, n, n), None, a, dev(B, cuda), b, Xd, 1e-12)
    assert rel_err(Xd.cpu().numpy(), Xo) < 1e-9


def test_cg_special_cases(cuda):
    rp, ci, val, n = grid_system(8)
    k = 2
    A = MatrixCSR(rp, ci, val, n, n)
    cg = ConjugateGradient(n, 50, k)
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    assert cg(A, None, 1.0, torch.zeros(n, k, dtype=torch.float64, device=cuda), -1.0, X, 1e-8)
    assert torch.all(X == -3.0)
    Bn = torch.ones(n, k, dtype=torch.float64, device=cuda)
    Bn[5, 1] = float("nan")
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    assert not cg(A, None, 1.0, Bn, 0.0, X, 1e-8)
    assert torch.all(X == 3.0)
    cg0 = ConjugateGradient(n, 0, k)
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    assert not cg0(A, None, 1.0, torch.ones(n, k, dtype=torch.float64, device=cuda), 0.0, X, 1e-8)
    assert torch.all(X == 0.0) and cg0.IterationCount() == 0


def test_cg_initial_guess(cuda):
    rp, ci, val, n = grid_system(10)
    k = 1
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    B = np.random.default_rng(2).uniform(-1, 1, (n, k))
    Xs = np.linalg.solve(Dm, B)
    cg = ConjugateGradient(n, 100, k)
    Xd = dev(Xs, cuda)
    assert cg(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 1e-8, use_initial_guess=True)
    assert cg.IterationCount() == 0
    Xd = dev(Xs + 1e-3, cuda)
    Xo = Xs + 1e-3
    ok_o, it_o, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 100, 1e-12, use_guess=True)
    assert cg(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 1e-12, use_initial_guess=True)
    assert abs(cg.IterationCount() - it_o) <= 1 and rel_err(Xd.cpu().numpy(), Xo) < 1e-9


def test_cg_fp32_path(cuda):
    rp, ci, val, n = grid_system(20, 1.0)
    k = 4
    B = gen.uniform(5, n * k).reshape(n, k)
    cg = ConjugateGradient(n, 500, k, dtype=torch.float32)
    Xd = torch.zeros(n, k, dtype=torch.float32, device=cuda)
    assert cg(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda, torch.float32), 0.0, Xd, 1e-5)
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    assert rel_err(Xd.cpu().numpy(), np.linalg.solve(Dm, B)) < 1e-4


def test_cg_3d_laplacian_medium(cuda):
    """3D 7-point Laplacian + 1e-2 shift (the configs[2] operator) at 32^3, 8 RHS."""
    g, k = 32, 8
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    B = gen.uniform(3003, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    ok_o, it_o, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 2000, 1e-10, threads=8)
    cg = ConjugateGradient(n, 2000, k)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    assert cg(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 1e-10)
    assert ok_o == 1 and abs(cg.IterationCount() - it_o) <= 2
    assert rel_err(Xd.cpu().numpy(), Xo) < 1e-7


@pytest.mark.parametrize("prec", [False, True])
def test_cg_callback_operators_match_builtin(cuda, fused_knob, prec):
    """tg_cg_solve_op with host callbacks (the reference's generic A / P callables) gives
    bit-identical results to the built-in CSR / diagonal path in its four-kernel form
    (the fused form sums p.u over another workgroup partition: same result to rounding,
    covered below)."""
    from tensors_amd.solvers import cg_solve_op, csr_callback
    fused_knob(0)
    rp, ci, val, n = grid_system(16, 0.2)
    k = 3
    B = gen.uniform(77, n * k).reshape(n, k)
    A = MatrixCSR(rp, ci, val, n, n)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
    D = MatrixCSR(np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), d, n, n)
    cg = ConjugateGradient(n, 500, k)
    X1 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    assert cg(A, dev(d, cuda) if prec else None, 1.0, dev(B, cuda), 0.0, X1, 1e-10)
    fA = csr_callback(A)
    fP = csr_callback(D) if prec else None
    X2 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, it, rr = cg_solve_op(n, k, 500, fA, fP, 1.0, dev(B, cuda), 0.0, X2, 1e-10)
    assert ok and it == cg.IterationCount()
    assert torch.equal(X1, X2)


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("k", [1, 2, 3, 8, 33, 64, 130, 256])
@pytest.mark.parametrize("prec", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cg_fused_iteration_matches_four_kernel_path(cuda, fused_knob, form, k, prec, dtype):
    """The fused CG iteration forms against the four-kernel iteration and the oracle:
    form 1 computes p.u in the SpMM epilogue, form 2 forms p = z + beta p, u = A p and p.u
    in one kernel and x, r and r.z in a second.  Every lane-group width and vector width of
    the fused kernel (k = 1 ... 256), FP64 / FP32, P = I / Jacobi, int64 and int32 row
    pointers."""
    rp, ci, val, n = grid_system(20, 0.1)
    if k % 2:
        rp = rp.astype(np.int32)
    B = gen.uniform(911 + k, n * k).reshape(n, k)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    A = MatrixCSR(rp, ci, val, n, n)
    out = {}
    for fused in (0, form):
        fused_knob(fused)
        cg = ConjugateGradient(n, 1000, k, dtype=dtype)
        Xd = torch.zeros(n, k, dtype=dtype, device=cuda)
        assert cg(A, None if d is None else dev(d, cuda, dtype), 1.0, dev(B, cuda, dtype), 0.0, Xd, tol)
        out[fused] = (Xd.double().cpu().numpy(), cg.IterationCount(), cg.RelativeResiduals())
    (X0, it0, rr0), (X1, it1, rr1) = out[0], out[form]
    assert abs(it0 - it1) <= 1
    assert np.all(rr1 <= tol)
    assert rel_err(X1, X0) < (1e-9 if dtype == torch.float64 else 2e-4)
    if dtype == torch.float64:
        Xo = np.zeros((n, k))
        ok_o, it_o, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 1000, tol, prec_diag=d)
        assert ok_o == 1 and abs(it1 - it_o) <= 1
        assert rel_err(X1, Xo) < 1e-8


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("k", [1, 8])
@pytest.mark.parametrize("prec", [False, True])
def test_cg_graph_blocks_equal_eager_launches(cuda, fused_knob, form, k, prec):
    """Knob cg_graph: after the first batch, blocks of 16 iterations run as one captured
    hipGraph.  The same kernels in the same order as eager launches, so the solution, the
    iteration count and the residuals are bitwise the same, including a solve that stops
    inside a block and one that runs into max_iter; and the mixed solver (host hook every
    16 iterations) likewise."""
    rp, ci, val, n = grid_system(40, 0.05)
    B = gen.uniform(977 + k, n * k).reshape(n, k)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    A = MatrixCSR(rp, ci, val, n, n)
    fused_knob(form)
    before = int(capi.lib().tg_tune_get(b"cg_graph"))
    out = {}
    try:
        for mode in (1, 0):
            capi.call("tg_tune_set", b"cg_graph", mode)
            res = []
            for max_iter, tol in ((2000, 1e-10), (45, 1e-14)):
                cg = ConjugateGradient(n, max_iter, k)
                Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
                ok = cg(A, None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, tol)
                res.append((ok, cg.IterationCount(), cg.RelativeResiduals().tolist(), Xd.cpu()))
            Xm = torch.zeros(n, k, dtype=torch.float64, device=cuda)
            mixed = cg_solve_mixed(A, dev(B, cuda), Xm, tol=1e-10, P=None if d is None else dev(d, cuda))
            res.append((mixed[:3], Xm.cpu()))
            out[mode] = res
    finally:
        capi.call("tg_tune_set", b"cg_graph", before)
    (a1, a2, am), (b1, b2, bm) = out[1], out[0]
    assert a1[0] and a1[1] > 40 and not a2[0] and a2[1] == 45
    for x, y in ((a1, b1), (a2, b2)):
        assert x[:3] == y[:3] and torch.equal(x[3], y[3])
    assert am[0] == bm[0] and torch.equal(am[1], bm[1])


def _cg_raw(A, n, k, dtype, d, B, max_iter, tol, ws_bytes, cuda, X0=None):
    """tg_cg_solve with a workspace of exactly ws_bytes bytes (X0: initial guess)."""
    import ctypes as C
    tg = capi.F64 if dtype == torch.float64 else capi.F32
    rp, ci, vals = A.device_arrays(dtype)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=cuda)
    X = torch.zeros(n, k, dtype=dtype, device=cuda) if X0 is None else X0.clone()
    ok, it = C.c_int32(0), C.c_int32(0)
    rr = np.zeros(k)
    capi.call("tg_cg_solve", n, A.NonzeroCount(), rp.data_ptr(), capi.I64 if rp.dtype == torch.int64 else capi.I32,
              ci.data_ptr(), vals.data_ptr(), tg, 0 if d is None else 1, None if d is None else d.data_ptr(), k,
              max_iter, 1.0, B.data_ptr(), k, 0.0, X.data_ptr(), k, tol, 0 if X0 is None else 1, ws.data_ptr(),
              ws_bytes, C.byref(ok), C.byref(it), rr.ctypes.data, None)
    torch.cuda.synchronize()
    return bool(ok.value), it.value, rr.tolist(), X.cpu()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("k", [1, 8])
@pytest.mark.parametrize("prec", [False, True])
def test_cg_deferred_x_update_is_bitwise(cuda, fused_knob, dtype, k, prec):
    """Knob cg_xdefer (form 1): x += alpha p of 1, 2 or 4 iterations in one pass from a
    ring of p.  The updates are the same fused multiply-adds in the same order, so the
    solution, iteration count and residuals are bitwise those of the per-iteration update:
    a converged solve (stopping inside a group of 4), one that runs into max_iter = 45, a
    workspace with only the two ring slots every solve accepts (depth 2 then), a solve from
    an initial guess, and the mixed solver, whose host hook folds x every 16 iterations."""
    rp, ci, val, n = grid_system(40, 0.05)
    B = gen.uniform(1313 + k, n * k).reshape(n, k)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    A = MatrixCSR(rp, ci, val, n, n)
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    vs = 8 if dtype == torch.float64 else 4
    full = int(capi.lib().tg_cg_workspace_bytes(n, k, capi.F64 if vs == 8 else capi.F32, 0 if d is None else 1))
    vec = (n * k * vs + 255) // 256 * 256
    dd = None if d is None else dev(d, cuda, dtype)
    Bd = dev(B, cuda, dtype)
    X0 = dev(gen.uniform(1414 + k, n * k).reshape(n, k), cuda, dtype)
    fused_knob(1)
    before = int(capi.lib().tg_tune_get(b"cg_xdefer"))
    out = {}
    try:
        for depth in (1, 2, 4):
            capi.call("tg_tune_set", b"cg_xdefer", depth)
            res = [_cg_raw(A, n, k, dtype, dd, Bd, mi, t, full, cuda) for mi, t in ((2000, tol), (45, 1e-14))]
            res.append(_cg_raw(A, n, k, dtype, dd, Bd, 2000, tol, full - 2 * vec, cuda))
            res.append(_cg_raw(A, n, k, dtype, dd, Bd, 2000, tol, full, cuda, X0=X0))  # x starts at a guess
            if dtype == torch.float64:
                Xm = torch.zeros(n, k, dtype=torch.float64, device=cuda)
                mixed = cg_solve_mixed(A, Bd, Xm, tol=1e-10, P=dd)
                res.append((mixed[0], mixed[1], mixed[2], Xm.cpu()))
            out[depth] = res
    finally:
        capi.call("tg_tune_set", b"cg_xdefer", before)
    conv, capped = out[1][0], out[1][1]
    assert conv[0] and conv[1] > 40
    assert not capped[0] and capped[1] == 45  # 45 = 11 groups of 4 and one pending update
    for depth in (2, 4):
        for a, b in zip(out[depth], out[1]):
            assert a[:3] == b[:3] and torch.equal(a[3], b[3])


@pytest.mark.parametrize("n,k", [(1, 1), (1, 4), (2, 3), (3, 8)])
@pytest.mark.parametrize("prec", [False, True])
def test_cg_tiny_systems(cuda, n, k, prec):
    """Systems of 1-3 rows (the Krylov space is used up after n iterations): the drop-in CG
    and the oracle take the same iterations to the same solution."""
    A = sp.diags([np.full(n - 1, -1.0), np.full(n, 2.4), np.full(n - 1, -1.0)], [-1, 0, 1], format="csr") \
        if n > 1 else sp.csr_matrix(np.array([[2.0]]))
    A.sort_indices()
    rp, ci, val = A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data
    B = gen.uniform(17 + n + k, n * k).reshape(n, k)
    d = 1.0 / A.diagonal() if prec else None
    cg = ConjugateGradient(n, 50, k)
    X = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok = cg(MatrixCSR(rp, ci, val, n, n), None if d is None else dev(d, cuda, torch.float64), 1.0,
            dev(B, cuda, torch.float64), 0.0, X, 1e-10)
    Xo = np.zeros((n, k))
    ok_o, it_o, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 50, 1e-10, prec_diag=d)
    assert bool(ok) == bool(ok_o) and abs(cg.IterationCount() - it_o) <= 1
    assert rel_err(X.cpu().numpy(), np.linalg.solve(A.toarray(), B)) < 1e-9
