"""Mixed-precision CG (FP32 inner sweeps + FP64 refinement) against the FP64 oracle:
the true FP64 residual reaches 1e-10 and the solution matches the oracle's FP64 CG."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import _capi as capi
from tensors_amd import generators as gen
from tensors_amd.solvers import cg_solve_mixed
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 4, 32])
@pytest.mark.parametrize("prec", [False, True])
def test_mixed_reaches_true_residual_and_matches_oracle(cuda, k, prec):
    g = 24
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    B = gen.uniform(3003 + k, n * k).reshape(n, k)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    A = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, outer, inner, rr = cg_solve_mixed(A, torch.from_numpy(B).to(cuda), Xd, tol=1e-10,
                                          P=None if d is None else torch.from_numpy(d).to(cuda))
    assert ok and outer >= 1 and np.all(rr <= 1e-10)
    X = Xd.cpu().numpy()
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n))
    true = np.linalg.norm(B - Dm @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(true <= 1.5e-10)
    Xo = np.zeros((n, k))
    oracle.cg(n, rp, ci, val, B, Xo, k, 5000, 1e-12, threads=8)
    assert rel_err(X, Xo) < 1e-7


def test_mixed_alpha_beta_and_special_cases(cuda):
    g = 10
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n, k = g ** 3, 2
    A = MatrixCSR(rp, ci, val, n, n)
    rng = np.random.default_rng(0)
    B = rng.uniform(-1, 1, (n, k))
    X0 = rng.uniform(-1, 1, (n, k))
    Xd = torch.from_numpy(X0.copy()).to(cuda)
    ok, _, _, _ = cg_solve_mixed(A, torch.from_numpy(B).to(cuda), Xd, tol=1e-11, a=-2.0, b=0.5)
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    assert ok and rel_err(Xd.cpu().numpy(), -2.0 * np.linalg.solve(Dm, B) + 0.5 * X0) < 1e-9
    Xd = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    ok, outer, inner, _ = cg_solve_mixed(A, torch.zeros(n, k, dtype=torch.float64, device=cuda), Xd, b=2.0)
    assert ok and outer == 0 and torch.all(Xd == 6.0)
    Bn = torch.ones(n, k, dtype=torch.float64, device=cuda)
    Bn[1, 1] = float("nan")
    Xd = torch.full((n, k), 3.0, dtype=torch.float64, device=cuda)
    ok, _, _, _ = cg_solve_mixed(A, Bn, Xd)
    assert not ok and torch.all(Xd == 3.0)


def test_replacement_keeps_the_krylov_process(cuda):
    """Residual replacement keeps p, so the FP32 iteration count stays near FP64 CG's
    (a restart per refinement step would cost roughly one extra half-solve each)."""
    from tensors_amd.solvers import ConjugateGradient

    g, k = 40, 4
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    A = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    B = torch.from_numpy(gen.uniform(3003, n * k).reshape(n, k)).to(cuda)
    X64 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    cg = ConjugateGradient(n, 5000, k)
    assert cg(A, None, 1.0, B, 0.0, X64, 1e-10)
    it64 = cg.IterationCount()
    Xm = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, reps, it32, rr = cg_solve_mixed(A, B, Xm, tol=1e-10)
    assert ok and np.all(rr <= 1e-10) and reps >= 2
    assert it32 <= 1.15 * it64 + 16, (it32, it64)
    # an initial guess near the solution needs only a few iterations
    Xg = Xm.clone() * (1 + 1e-6)
    ok, _, it_g, rr = cg_solve_mixed(A, B, Xg, tol=1e-10, use_initial_guess=True)
    assert ok and it_g < it32 and np.all(rr <= 1e-10)


@pytest.fixture
def fused_knob():
    """Restores the process-wide cg_fused knob after the test."""
    before = int(capi.lib().tg_tune_get(b"cg_fused"))
    yield lambda v: capi.call("tg_tune_set", b"cg_fused", int(v))
    capi.call("tg_tune_set", b"cg_fused", before)


@pytest.mark.parametrize("n,k", [(1, 1), (2, 8), (3, 3), (5, 32)])
@pytest.mark.parametrize("prec", [False, True])
@pytest.mark.parametrize("form", [3, 2], ids=["default", "fused_direction"])
def test_mixed_tiny_systems(cuda, fused_knob, n, k, prec, form):
    """Systems smaller than a handful of CG iterations: the FP32 iteration uses up its Krylov
    space at once, and the solve must still reach the FP64 true residual -- in the default
    iteration and in the fused-direction form, whose restart after a replacement comes from
    beta = 0 instead of a reset p."""
    fused_knob(form)
    A = sp.diags([np.full(n - 1, -1.0), np.full(n, 2.3), np.full(n - 1, -1.0)], [-1, 0, 1], format="csr") \
        if n > 1 else sp.csr_matrix(np.array([[2.0]]))
    A.sort_indices()
    B = gen.uniform(31 + n, n * k).reshape(n, k)
    M = MatrixCSR(A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, n, n)
    P = torch.from_numpy(1.0 / A.diagonal()).to(cuda) if prec else None
    X = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, outer, inner, rr = cg_solve_mixed(M, torch.from_numpy(B).to(cuda), X, tol=1e-10, P=P)
    assert ok, (outer, inner, rr)
    Xh = X.cpu().numpy()
    assert np.max(np.linalg.norm(B - A @ Xh, axis=0) / np.linalg.norm(B, axis=0)) <= 1.5e-10
