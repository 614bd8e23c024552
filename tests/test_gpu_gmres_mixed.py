"""Mixed-precision GMRES (FP32 Arnoldi cycles, FP64 residual / update / stop test at every
restart, FP64 Hessenberg, Givens and least squares) on the nonsymmetric upwind
convection-diffusion operator: the true FP64 residual reaches the tolerance, including
tolerances below FP32 precision, the solution matches a dense FP64 solve and the oracle's
FP64 GMRES, and the restart count stays close to FP64 GMRES(m)'s."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.solvers import GMRES, gmres_solve_mixed
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu


def problem(g, peclet):
    rp, ci, val = gen.convection_diffusion_2d(g, peclet)
    n = len(rp) - 1
    A = sp.csr_matrix((val, ci, rp), shape=(n, n))
    return rp, ci, val, n, A


@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("tol", [1e-8, 1e-11])
@pytest.mark.parametrize("side,prec", [(0, False), (0, True), (1, True)], ids=["plain", "left_jacobi", "right_jacobi"])
def test_mixed_gmres_reaches_fp64_residual(cuda, k, tol, side, prec):
    rp, ci, val, n, A = problem(48, 10.0)
    d = 1.0 / A.diagonal() if prec else None
    B = gen.uniform(5005 + k, n * k).reshape(n, k)
    M = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    Pd = torch.from_numpy(d).to(cuda) if prec else None
    X = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, restarts, steps, rr = gmres_solve_mixed(M, torch.from_numpy(B).to(cuda), X, tol=tol, restart=30,
                                                max_restarts=200, P=Pd, side=side)
    print(f"mixed GMRES side {side} prec {prec} k {k} tol {tol}: restarts {restarts} steps {steps} {rr}")
    assert ok and np.all(rr <= tol)
    Xh = X.cpu().numpy()
    R = B - A @ Xh
    if side == 0 and prec:
        true = np.linalg.norm(d[:, None] * R, axis=0) / np.linalg.norm(d[:, None] * B, axis=0)
    else:
        true = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    # the device's FP64 residual and the host's agree to rounding of b - A x (in units of ||P b||)
    assert np.all(true <= 1.5 * tol) and np.allclose(true, rr, rtol=1e-6, atol=1e-14)
    Xs = np.linalg.solve(A.toarray(), B)
    cond_bound = 1e4 * tol  # the operator's condition number is below 1e4 at this size
    assert rel_err(Xh, Xs) < cond_bound
    # the oracle's FP64 GMRES(30) at the same tolerance: same solution within the bound,
    # and the mixed form needs at most a few more restarts
    if tol >= 1e-8:
        Xo = np.zeros((n, k))
        ok_o, it_o, rs_o, _ = oracle.gmres(n, rp, ci, val, B, Xo, k, 30, tol, 200, prec_diag=d, side=side)
        assert ok_o and rel_err(Xh, Xo) < cond_bound
        assert restarts <= rs_o + 3, (restarts, rs_o)


def test_mixed_gmres_restarts_against_fp64_gpu(cuda):
    """On a 200^2 grid (Peclet 30) the mixed solve restarts about as often as the FP64 GPU
    GMRES(30) and both reach 1e-8."""
    rp, ci, val, n, A = problem(200, 30.0)
    B = gen.uniform(77, n).reshape(n, 1)
    M = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    Bd = torch.from_numpy(B).to(cuda)
    X64 = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
    gm = GMRES(n, 30, 1)
    assert gm(M, None, 1.0, Bd, 0.0, X64, 1e-8, 500)
    Xm = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
    ok, restarts, steps, rr = gmres_solve_mixed(M, Bd, Xm, tol=1e-8, restart=30, max_restarts=500)
    print(f"FP64 restarts {gm.RestartCount()}, mixed restarts {restarts} ({steps} FP32 steps)")
    assert ok and rr[0] <= 1e-8
    assert restarts <= gm.RestartCount() + 3
    true = np.linalg.norm(B - A @ Xm.cpu().numpy()) / np.linalg.norm(B)
    assert true <= 1.5e-8


def test_mixed_gmres_special_cases(cuda):
    rp, ci, val, n, A = problem(12, 3.0)
    M = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    k = 2
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    ok, *_ = gmres_solve_mixed(M, torch.zeros(n, k, dtype=torch.float64, device=cuda), X, b=0.0)
    assert ok and torch.all(X == 0.0)  # zero right-hand side, b = 0: X zeroed (GMRES.hpp:206-229)
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    ok, *_ = gmres_solve_mixed(M, torch.zeros(n, k, dtype=torch.float64, device=cuda), X, b=-1.0)
    assert ok and torch.all(X == -3.0)
    Bn = torch.ones(n, k, dtype=torch.float64, device=cuda)
    Bn[3, 1] = float("nan")
    X = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    ok, *_ = gmres_solve_mixed(M, Bn, X)
    assert not ok and torch.all(X == 3.0)  # NaN: nothing written
    B = torch.from_numpy(gen.uniform(3, n * k).reshape(n, k)).to(cuda)
    X0 = torch.from_numpy(gen.uniform(4, n * k).reshape(n, k)).to(cuda)
    X = X0.clone()
    ok, *_ = gmres_solve_mixed(M, B, X, tol=1e-10, a=2.0, b=0.5)
    Xs = np.linalg.solve(A.toarray(), B.cpu().numpy())
    assert ok and rel_err(X.cpu().numpy(), 2.0 * Xs + 0.5 * X0.cpu().numpy()) < 1e-7


@pytest.mark.parametrize("g,side,prec", [(60, 0, False), (60, 0, True), (40, 1, True), (800, 0, False)],
                         ids=["plain", "left_jacobi", "right_jacobi", "large"])
def test_mixed_gmres_split_dots_agree(cuda, g, side, prec):
    """Knob gmres_split_dots: the FP32 dots pass of the 32-vector bucket with the u-side and
    v-side products on alternate lanes (one row per iteration below 2^19 rows, two above)
    against both sides on every lane.  Only the summation order differs, so both reach the
    tolerance on the true FP64 residual in the same number of restarts (within one) and
    the solutions agree far below the tolerance's scale."""
    from tensors_amd import _capi as capi

    rp, ci, val, n, A = problem(g, 20.0)
    d = 1.0 / A.diagonal() if prec else None
    B = gen.uniform(6006 + g, n).reshape(n, 1)
    M = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    Pd = torch.from_numpy(d).to(cuda) if prec else None
    before = int(capi.lib().tg_tune_get(b"gmres_split_dots"))
    out = {}
    try:
        for split in (1, 0):
            capi.call("tg_tune_set", b"gmres_split_dots", split)
            X = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
            ok, restarts, steps, rr = gmres_solve_mixed(M, torch.from_numpy(B).to(cuda), X, tol=1e-8, restart=30,
                                                        max_restarts=300, P=Pd, side=side)
            out[split] = (ok, restarts, rr, X.cpu().numpy())
    finally:
        capi.call("tg_tune_set", b"gmres_split_dots", before)
    (ok1, r1, rr1, X1), (ok0, r0, rr0, X0) = out[1], out[0]
    print(f"split dots g {g}: restarts {r1} / {r0}, residuals {rr1} / {rr0}")
    assert ok1 and ok0 and np.all(rr1 <= 1e-8) and np.all(rr0 <= 1e-8)
    assert r1 >= 1 and abs(r1 - r0) <= 1  # the 32-vector bucket ran (a whole cycle of 30 steps)
    assert rel_err(X1, X0) < 1e-6
