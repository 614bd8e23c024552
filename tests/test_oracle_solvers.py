"""Pins the oracle's CG, GMRES and back substitution (oracle.cpp) against what the
mathematics fixes: dense solves (numpy), finite termination of CG in at most
#distinct-eigenvalue steps on diagonal systems, GMRES(n) exactness, the invariant
|beta_iter| = ||P(b - A x)|| at convergence, numpy's triangular solve, and the
special cases of ConjugateGradient.hpp:147-211 and GMRES.hpp:193-229."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import rel_err


def spd_matrix(n, seed):
    rng = np.random.default_rng(seed)
    Bm = sp.random(n, n, density=0.1, random_state=rng)
    A = (Bm @ Bm.T + sp.eye(n) * 2.0).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.copy(), A.toarray()


@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("prec", [False, True])
def test_cg_matches_dense_solve(k, prec):
    n = 60
    rp, ci, val, D = spd_matrix(n, 1)
    rng = np.random.default_rng(2)
    B = rng.uniform(-1, 1, (n, k))
    X = np.zeros((n, k))
    d = 1.0 / np.diag(D) if prec else None
    ok, it, rr = oracle.cg(n, rp, ci, val, B, X, k, 500, 1e-12, prec_diag=d)
    assert ok == 1 and 0 < it <= 500
    assert rel_err(X, np.linalg.solve(D, B)) < 1e-9
    assert np.all(rr <= 1e-12)


@pytest.mark.parametrize("a,b", [(1.0, 0.0), (-1.0, 1.0), (np.pi, -1.0), (2.0, 1.0)])
def test_cg_alpha_beta_combination(a, b):
    n, k = 40, 2
    rp, ci, val, D = spd_matrix(n, 3)
    rng = np.random.default_rng(4)
    B = rng.uniform(-1, 1, (n, k))
    X0 = rng.uniform(-1, 1, (n, k))
    X = X0.copy()
    ok, _, _ = oracle.cg(n, rp, ci, val, B, X, k, 500, 1e-13, a=a, b=b)
    assert ok == 1
    assert rel_err(X, a * np.linalg.solve(D, B) + b * X0) < 1e-9


def test_cg_finite_termination_on_diagonal():
    """Diagonal A with 4 distinct eigenvalues: CG is exact after 4 steps."""
    n = 50
    diag = np.array([1.0, 2.0, 5.0, 9.0])[np.arange(n) % 4]
    rp = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int32)
    B = np.random.default_rng(0).uniform(-1, 1, (n, 1))
    X = np.zeros((n, 1))
    ok, it, _ = oracle.cg(n, rp, ci, diag, B, X, 1, 100, 1e-14)
    assert ok == 1 and it <= 5
    assert rel_err(X[:, 0], B[:, 0] / diag) < 1e-12


def test_cg_special_cases():
    n, k = 20, 2
    rp, ci, val, _ = spd_matrix(n, 5)
    X0 = np.full((n, k), 3.0)
    X = X0.copy()
    ok, it, _ = oracle.cg(n, rp, ci, val, np.ones((n, k)), X, k, 10, 1e-8, a=0.0, b=2.0)
    assert ok == 1 and it == 0 and np.all(X == 6.0)          # a == 0 -> X = b X
    X = X0.copy()
    ok, it, _ = oracle.cg(n, rp, ci, val, np.zeros((n, k)), X, k, 10, 1e-8, a=1.0, b=-1.0)
    assert ok == 1 and np.all(X == -3.0)                      # B == 0 -> X = b X
    Bn = np.ones((n, k))
    Bn[3, 1] = np.nan
    X = X0.copy()
    ok, _, _ = oracle.cg(n, rp, ci, val, Bn, X, k, 10, 1e-8)
    assert ok == 0 and np.all(X == 3.0)                       # NaN -> nothing written
    X = X0.copy()
    ok, it, _ = oracle.cg(n, rp, ci, val, np.ones((n, k)), X, k, 0, 1e-8, b=0.0)
    assert ok == 0 and it == 0 and np.all(X == 0.0)           # max_iter 0: x = 0


def test_cg_initial_guess_exact_solution_stops_at_zero_iterations():
    n, k = 30, 1
    rp, ci, val, D = spd_matrix(n, 6)
    B = np.random.default_rng(1).uniform(-1, 1, (n, k))
    X = np.linalg.solve(D, B)
    ok, it, _ = oracle.cg(n, rp, ci, val, B, X, k, 50, 1e-8, use_guess=True)
    assert ok == 1 and it == 0


def nonsym_matrix(n, seed):
    rng = np.random.default_rng(seed)
    A = (sp.random(n, n, density=0.15, random_state=rng) + sp.eye(n) * 4.0).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.copy(), A.toarray()


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("k", [1, 4])
def test_gmres_matches_dense_solve(side, k):
    n = 50
    rp, ci, val, D = nonsym_matrix(n, 7)
    B = np.random.default_rng(8).uniform(-1, 1, (n, k))
    X = np.zeros((n, k))
    d = 1.0 / np.diag(D)
    ok, it, rs, rr = oracle.gmres(n, rp, ci, val, B, X, k, 30, 1e-12, 20, prec_diag=d, side=side)
    assert ok == 1
    assert rel_err(X, np.linalg.solve(D, B)) < 1e-9


def test_gmres_full_krylov_space_is_exact():
    """GMRES(n) without restarts solves an n x n system in at most n steps."""
    n = 12
    rp, ci, val, D = nonsym_matrix(n, 9)
    B = np.random.default_rng(1).uniform(-1, 1, (n, 1))
    X = np.zeros((n, 1))
    ok, it, rs, _ = oracle.gmres(n, rp, ci, val, B, X, 1, n, 1e-13, 0)
    assert ok == 1 and it <= n and rs == 0
    assert rel_err(X, np.linalg.solve(D, B)) < 1e-10


def test_gmres_alpha_beta_and_restarts():
    n, k = 80, 2
    rp, ci, val, D = nonsym_matrix(n, 10)
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (n, k))
    X0 = rng.uniform(-1, 1, (n, k))
    X = X0.copy()
    ok, it, rs, rr = oracle.gmres(n, rp, ci, val, B, X, k, 5, 1e-12, 200, a=-2.0, b=0.5)
    assert ok == 1 and rs > 0
    assert rel_err(X, -2.0 * np.linalg.solve(D, B) + 0.5 * X0) < 1e-9
    # |beta_iter| / ||b|| reported is the true relative residual at the final cycle
    assert np.all(rr <= 1e-12)


def test_gmres_special_cases():
    n, k = 20, 2
    rp, ci, val, _ = nonsym_matrix(n, 11)
    X = np.full((n, k), 3.0)
    ok, it, rs, _ = oracle.gmres(n, rp, ci, val, np.zeros((n, k)), X, k, 10, 1e-8, 5, b=0.0)
    assert ok == 1 and np.all(X == 0.0)
    X = np.full((n, k), 3.0)
    ok, _, _, _ = oracle.gmres(n, rp, ci, val, np.zeros((n, k)), X, k, 10, 1e-8, 5, b=2.0)
    assert ok == 1 and np.all(X == 6.0)
    Bn = np.ones((n, k))
    Bn[0, 0] = np.nan
    X = np.full((n, k), 3.0)
    ok, _, _, _ = oracle.gmres(n, rp, ci, val, Bn, X, k, 10, 1e-8, 5)
    assert ok == 0 and np.all(X == 3.0)


def test_trsv_upper_matches_numpy():
    rng = np.random.default_rng(4)
    for k in (1, 5, 30):
        H = np.triu(rng.uniform(-1, 1, (k, k))) + np.eye(k) * 3.0
        b = rng.uniform(-1, 1, k)
        x = b.copy()
        oracle.lib().orc_trsv_upper(k, oracle.p(H), k, oracle.p(x))
        assert rel_err(x, np.linalg.solve(H, b)) < 1e-13


def test_solver_wrappers_accept_int32_row_pointers():
    """The CG / GMRES wrappers convert rp to the C signature's int64 (an int32 array used to
    be read as int64 out of bounds)."""
    rp, ci, val = oracle.grid_laplacian(12, 0.3)
    n, k = 144, 3
    B = np.random.default_rng(5).uniform(-1, 1, (n, k))
    X64, X32 = np.zeros((n, k)), np.zeros((n, k))
    r64 = oracle.cg(n, rp, ci, val, B, X64, k, 500, 1e-10)
    r32 = oracle.cg(n, rp.astype(np.int32), ci, val, B, X32, k, 500, 1e-10)
    assert r64[:2] == r32[:2] and np.array_equal(X64, X32)
    G64, G32 = np.zeros((n, k)), np.zeros((n, k))
    g64 = oracle.gmres(n, rp, ci, val, B, G64, k, 30, 1e-10, 50)
    g32 = oracle.gmres(n, rp.astype(np.int32), ci, val, B, G32, k, 30, 1e-10, 50)
    assert g64[:3] == g32[:3] and np.array_equal(G64, G32)
