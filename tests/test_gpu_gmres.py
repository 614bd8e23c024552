"""GPU parity of tg_gmres_solve (GMRES(m) behind GMRES::operator()) against the oracle's
restatement of GMRES.hpp (left/right preconditioning, restarts, alpha/beta) and dense
solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.solvers import GMRES
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu


def dev(a, d, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dtype)


def cd_system(g, pe):
    rp, ci, val = gen.convection_diffusion_2d(g, pe)
    return rp, ci, val, g * g


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("k", [1, 4])
@pytest.mark.parametrize("prec", [False, True])
def test_gmres_matches_oracle(cuda, side, k, prec):
    rp, ci, val, n = cd_system(20, 3.0)
    B = gen.uniform(5005 + k, n * k).reshape(n, k)
    d = None
    if prec:
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
    Xo = np.zeros((n, k))
    ok_o, it_o, rs_o, rr_o = oracle.gmres(n, rp, ci, val, B, Xo, k, 30, 1e-10, 100, prec_diag=d, side=side)
    g = GMRES(n, 30, k, side=side)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok = g(MatrixCSR(rp, ci, val, n, n), None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, 1e-10,
           100)
    assert ok and ok_o == 1
    assert abs(g.RestartCount() - rs_o) <= 1
    assert rel_err(Xd.cpu().numpy(), Xo) < 1e-7
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    assert rel_err(Xd.cpu().numpy(), np.linalg.solve(Dm, B)) < 1e-7


@pytest.mark.parametrize("a,b", [(-1.0, 0.0), (2.0, 1.0), (np.pi, -0.5)])
def test_gmres_alpha_beta(cuda, a, b):
    rp, ci, val, n = cd_system(12, 10.0)
    k = 2
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (n, k))
    X0 = rng.uniform(-1, 1, (n, k))
    Xo = X0.copy()
    oracle.gmres(n, rp, ci, val, B, Xo, k, 20, 1e-12, 50, a=a, b=b)
    g = GMRES(n, 20, k)
    Xd = dev(X0, cuda)
    assert g(MatrixCSR(rp, ci, val, n, n), None, a, dev(B, cuda), b, Xd, 1e-12, 50)
    assert rel_err(Xd.cpu().numpy(), Xo) < 1e-8


def test_gmres_full_space_and_special_cases(cuda):
    rp, ci, val, n = cd_system(4, 2.0)  # n = 16: GMRES(16) is exact
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    B = np.random.default_rng(0).uniform(-1, 1, (n, 1))
    g = GMRES(n, n, 1)
    Xd = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
    A = MatrixCSR(rp, ci, val, n, n)
    assert g(A, None, 1.0, dev(B, cuda), 0.0, Xd, 1e-13, 0)
    assert g.IterationCount() <= n and g.RestartCount() == 0
    assert rel_err(Xd.cpu().numpy(), np.linalg.solve(Dm, B)) < 1e-10
    X = torch.full((n, 1), 3.0, dtype=torch.float64, device=cuda)
    assert g(A, None, 1.0, torch.zeros(n, 1, dtype=torch.float64, device=cuda), 0.0, X, 1e-8, 5)
    assert torch.all(X == 0.0)
    X = torch.full((n, 1), 3.0, dtype=torch.float64, device=cuda)
    assert g(A, None, 1.0, torch.zeros(n, 1, dtype=torch.float64, device=cuda), 2.0, X, 1e-8, 5)
    assert torch.all(X == 6.0)
    Bn = torch.ones(n, 1, dtype=torch.float64, device=cuda)
    Bn[2, 0] = float("nan")
    X = torch.full((n, 1), 3.0, dtype=torch.float64, device=cuda)
    assert not g(A, None, 1.0, Bn, 0.0, X, 1e-8, 5)
    assert torch.all(X == 3.0)


def test_gmres_initial_guess_and_fp32(cuda):
    rp, ci, val, n = cd_system(16, 5.0)
    Dm = sp.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    B = np.random.default_rng(1).uniform(-1, 1, (n, 1))
    Xs = np.linalg.solve(Dm, B)
    g = GMRES(n, 30, 1)
    Xd = dev(Xs, cuda)
    assert g(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 1e-8, 10, use_initial_guess=True)
    assert g.IterationCount() == 0
    g32 = GMRES(n, 30, 1, dtype=torch.float32)
    X32 = torch.zeros(n, 1, dtype=torch.float32, device=cuda)
    assert g32(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda, torch.float32), 0.0, X32, 1e-5, 50)
    assert rel_err(X32.cpu().numpy(), Xs) < 1e-4


def test_gmres_medium_convection_diffusion(cuda):
    """The serving-mix operator at g = 100 (n = 1e4), Peclet 30, GMRES(30) to 1e-8."""
    rp, ci, val, n = cd_system(100, 30.0)
    B = gen.uniform(5005, n).reshape(n, 1)
    Xo = np.zeros((n, 1))
    ok_o, it_o, rs_o, _ = oracle.gmres(n, rp, ci, val, B, Xo, 1, 30, 1e-8, 100)
    g = GMRES(n, 30, 1)
    Xd = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
    ok = g(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 1e-8, 100)
    assert ok == bool(ok_o)
    assert abs(g.RestartCount() - rs_o) <= 1
    assert rel_err(Xd.cpu().numpy(), Xo) < 1e-6


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("prec", [False, True])
def test_gmres_callback_operators_match_builtin(cuda, side, prec):
    """tg_gmres_solve_op with host callbacks for A and P agrees with the built-in path."""
    from tensors_amd.solvers import csr_callback, gmres_solve_op
    rp, ci, val, n = cd_system(14, 4.0)
    k = 2
    B = gen.uniform(91, n * k).reshape(n, k)
    A = MatrixCSR(rp, ci, val, n, n)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
    D = MatrixCSR(np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), d, n, n)
    g = GMRES(n, 25, k, side=side)
    X1 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    assert g(A, dev(d, cuda) if prec else None, 1.0, dev(B, cuda), 0.0, X1, 1e-11, 50)
    X2 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok, it, rs, rr = gmres_solve_op(n, k, 25, csr_callback(A), csr_callback(D) if prec else None, 1.0,
                                    dev(B, cuda), 0.0, X2, 1e-11, 50, side=side)
    assert ok and rs == g.RestartCount() and it == g.IterationCount()
    assert rel_err(X2.cpu().numpy(), X1.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("prec", [False, True])
@pytest.mark.parametrize("k", [1, 3])
def test_gmres_cycle_graph_equals_eager(cuda, side, prec, k):
    """Cycles after the first run as one captured hipGraph (knob gmres_graph): the same
    kernels with the same arguments, so the solution, iteration and restart counts are
    bitwise those of eager launches, over many restarts."""
    from tensors_amd import _capi as capi

    rp, ci, val, n = cd_system(60, 20.0)
    B = gen.uniform(77 + k, n * k).reshape(n, k)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    A = MatrixCSR(rp, ci, val, n, n)
    out = {}
    before = capi.lib().tg_tune_get(b"gmres_graph")
    try:
        for mode in (1, 0):
            capi.call("tg_tune_set", b"gmres_graph", mode)
            g = GMRES(n, 10, k, side=side)
            X = torch.zeros(n, k, dtype=torch.float64, device=cuda)
            ok = g(A, None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, X, 1e-9, 400)
            out[mode] = (ok, g.iter, g.restarts, X.cpu().numpy())
    finally:
        capi.call("tg_tune_set", b"gmres_graph", before)
    assert out[1][0] and out[1][2] >= 3, out[1][:3]
    assert out[1][:3] == out[0][:3]
    assert np.array_equal(out[1][3], out[0][3])
    S = sp.csr_matrix((val, ci, rp), shape=(n, n))
    assert np.max(np.linalg.norm(B - S @ out[1][3], axis=0) / np.linalg.norm(B, axis=0)) < 1e-7


@pytest.mark.parametrize("n", [1, 2, 4])
@pytest.mark.parametrize("side", [0, 1])
def test_gmres_tiny_systems(cuda, n, side):
    """GMRES(30) on systems of 1-4 rows (the Krylov space is used up after n steps, before the
    restart length): the dense solution, in at most n steps, without restarts."""
    A = sp.diags([np.full(n - 1, -1.3), np.full(n, 3.0), np.full(n - 1, -0.7)], [-1, 0, 1], format="csr") \
        if n > 1 else sp.csr_matrix(np.array([[3.0]]))
    A.sort_indices()
    rp, ci, val = A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data
    B = gen.uniform(23 + n, n * 2).reshape(n, 2)
    g = GMRES(n, 30, 2, side=side)
    X = torch.zeros(n, 2, dtype=torch.float64, device=cuda)
    assert g(MatrixCSR(rp, ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, X, 1e-10, 5)
    assert g.RestartCount() == 0 and g.IterationCount() <= n
    assert rel_err(X.cpu().numpy(), np.linalg.solve(A.toarray(), B)) < 1e-9


@pytest.mark.parametrize("side,prec", [(0, False), (0, True), (1, True)], ids=["plain", "left_jacobi", "right_jacobi"])
def test_gmres_split_dots_agree(cuda, side, prec):
    """Knob gmres_split_dots in FP64 (one right-hand side, more than 16 basis vectors,
    below 2^19 rows): the dots pass with its u-side and v-side products on alternate
    lanes against both on every lane.  Only the summation order differs: same restarts,
    solutions within rounding of each other and of the oracle's GMRES."""
    from tensors_amd import _capi as capi

    rp, ci, val, n = cd_system(50, 20.0)
    B = gen.uniform(7007, n).reshape(n, 1)
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    A = MatrixCSR(rp, ci, val, n, n)
    before = int(capi.lib().tg_tune_get(b"gmres_split_dots"))
    out = {}
    try:
        for split in (1, 0):
            capi.call("tg_tune_set", b"gmres_split_dots", split)
            g = GMRES(n, 30, 1, side=side)
            Xd = torch.zeros(n, 1, dtype=torch.float64, device=cuda)
            ok = g(A, None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, 1e-10, 200)
            out[split] = (ok, g.RestartCount(), Xd.cpu().numpy())
    finally:
        capi.call("tg_tune_set", b"gmres_split_dots", before)
    (ok1, r1, X1), (ok0, r0, X0) = out[1], out[0]
    assert ok1 and ok0 and r1 >= 1 and abs(r1 - r0) <= 1
    assert rel_err(X1, X0) < 1e-9
    Xo = np.zeros((n, 1))
    ok_o, it_o, rs_o, _ = oracle.gmres(n, rp, ci, val, B, Xo, 1, 30, 1e-10, 200, prec_diag=d, side=side)
    assert ok_o == 1 and abs(r1 - rs_o) <= 1 and rel_err(X1, Xo) < 1e-7
