"""Pins the oracle's SpMV/SpMM (oracle.cpp: spmv/spmm/multiply_dense) against what the
mathematics fixes: dense brute force (the Dot_True triple loop of
Unit_Test_SpMM/main.cpp:19-60, done in numpy), exact integer arithmetic, the alpha/beta
flag table of SpMM.hpp:28-131 / SpMV.hpp:14-29, empty rows (SpMM.hpp:395-410), the
binary a == nullptr case and the reference's no-op error cases."""
import itertools

import numpy as np
import pytest

import oracle
from helpers import dense_of, random_csr_np, rel_err, spmm_bound

ALPHAS = [-1.0, 0.0, 1.0, 2.0, np.pi]      # Test_SpMM.nb:144-172 / Test_CG.nb grids
BETAS = [-1.0, 0.0, 1.0, 2.0, np.pi]


def brute(D, alpha, X, beta, Y0):
    if alpha == 0.0:
        return Y0 * beta if beta not in (0.0, 1.0) else (np.zeros_like(Y0) if beta == 0.0 else Y0.copy())
    out = alpha * (D @ X)
    if beta != 0.0:
        out = out + beta * Y0
    return out


@pytest.mark.parametrize("nrhs,NRHS", [(1, 0), (1, 1), (3, 0), (4, 4), (16, 16), (7, 0)])
@pytest.mark.parametrize("alpha,beta", list(itertools.product(ALPHAS, BETAS)))
def test_spmm_vs_dense(nrhs, NRHS, alpha, beta):
    m, n = 37, 29
    rp, ci, val = random_csr_np(m, n, 0.2, seed=nrhs * 7 + 1, empty_rows=0.15)
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (n, nrhs))
    Y0 = rng.uniform(-1, 1, (m, nrhs))
    Y = Y0.copy()
    st = oracle.spmm(rp, ci, val, m, n, alpha, X, nrhs, beta, Y, nrhs, nrhs, NRHS)
    assert st == 0
    ref = brute(dense_of(m, n, rp, ci, val), alpha, X, beta, Y0)
    bound = spmm_bound(m, n, rp, ci, val, alpha, X, beta, Y0, 2.0 ** -53)
    assert np.all(np.abs(Y - ref) <= bound)


def test_exact_integer_arithmetic():
    """Small integers: every summation order is exact, so the result is pinned to the bit."""
    m, n, k = 50, 40, 5
    rng = np.random.default_rng(3)
    rp, ci, _ = random_csr_np(m, n, 0.3, seed=11)
    val = rng.integers(-4, 5, len(ci)).astype(np.float64)
    X = rng.integers(-8, 9, (n, k)).astype(np.float64)
    Y0 = rng.integers(-8, 9, (m, k)).astype(np.float64)
    Y = Y0.copy()
    assert oracle.spmm(rp, ci, val, m, n, 3.0, X, k, -2.0, Y, k, k) == 0
    ref = 3.0 * (dense_of(m, n, rp, ci, val) @ X) - 2.0 * Y0
    assert np.array_equal(Y, ref)


def test_beta_zero_ignores_nan_and_empty_rows():
    m, n, k = 20, 20, 3
    rp, ci, val = random_csr_np(m, n, 0.2, seed=2, empty_rows=0.3)
    X = np.ones((n, k))
    Y = np.full((m, k), np.nan)
    assert oracle.spmm(rp, ci, val, m, n, 1.0, X, k, 0.0, Y, k, k) == 0
    assert np.all(np.isfinite(Y))
    empty = np.diff(rp) == 0
    assert empty.any() and np.all(Y[empty] == 0.0)


def test_beta_one_leaves_empty_rows_untouched():
    m, n, k = 20, 20, 2
    rp, ci, val = random_csr_np(m, n, 0.2, seed=4, empty_rows=0.4)
    X = np.ones((n, k))
    Y0 = np.full((m, k), np.inf)
    Y = Y0.copy()
    assert oracle.spmm(rp, ci, val, m, n, 2.0, X, k, 1.0, Y, k, k) == 0
    empty = np.diff(rp) == 0
    assert np.all(np.isinf(Y[empty]))


def test_alpha_zero_scales_only():
    m, n, k = 10, 10, 4
    rp, ci, val = random_csr_np(m, n, 0.3, seed=6)
    X = np.full((n, k), np.nan)  # never read
    Y0 = np.arange(m * k, dtype=np.float64).reshape(m, k)
    for beta, want in [(0.0, np.zeros_like(Y0)), (1.0, Y0), (2.5, 2.5 * Y0)]:
        Y = Y0.copy()
        assert oracle.spmm(rp, ci, val, m, n, 0.0, X, k, beta, Y, k, k) == 0
        assert np.array_equal(Y, want)


def test_binary_matrix():
    m, n, k = 30, 25, 3
    rp, ci, _ = random_csr_np(m, n, 0.2, seed=8)
    rng = np.random.default_rng(1)
    X = rng.integers(-5, 6, (n, k)).astype(np.float64)
    Y = np.zeros((m, k))
    assert oracle.spmm(rp, ci, None, m, n, 1.0, X, k, 0.0, Y, k, k) == 0
    assert np.array_equal(Y, dense_of(m, n, rp, ci, None) @ X)


def test_error_cases_are_noops():
    m = n = 8
    rp, ci, val = random_csr_np(m, n, 0.5, seed=9)
    X = np.ones((n, 2))
    Y = X  # alias
    snap = Y.copy()
    assert oracle.spmm(rp, ci, val, m, n, 1.0, X, 2, 0.0, Y, 2, 2) == -2
    assert np.array_equal(Y, snap)
    Y2 = np.full((m, 2), 7.0)
    assert oracle.spmm(rp, ci, val, m, n, 1.0, X, 2, 0.0, Y2, 2, 2, NRHS=4) == -3
    assert np.all(Y2 == 7.0)


def test_strided_blocks_and_f32():
    m, n, k, ld = 41, 33, 5, 9
    rp, ci, val = random_csr_np(m, n, 0.25, seed=12)
    rng = np.random.default_rng(2)
    Xf = rng.uniform(-1, 1, (n, ld))
    Yf = rng.uniform(-1, 1, (m, ld))
    Y0 = Yf.copy()
    assert oracle.spmm(rp, ci, val, m, n, 1.5, Xf, ld, -0.5, Yf, ld, k) == 0
    ref = 1.5 * dense_of(m, n, rp, ci, val) @ Xf[:, :k] - 0.5 * Y0[:, :k]
    assert rel_err(Yf[:, :k], ref) < 1e-14
    assert np.array_equal(Yf[:, k:], Y0[:, k:])  # columns beyond nrhs untouched
    X32 = Xf.astype(np.float32)
    Y32 = np.zeros((m, ld), np.float32)
    assert oracle.spmm(rp, ci, val.astype(np.float32), m, n, 1.0, X32, ld, 0.0, Y32, ld, k) == 0
    assert rel_err(Y32[:, :k], dense_of(m, n, rp, ci, val) @ Xf[:, :k]) < 1e-5


def test_job_pointers():
    L = oracle.lib()
    for n_, T in [(10, 3), (7, 7), (100, 8), (5, 8)]:
        jp = [L.orc_job_pointer(n_, T, t) for t in range(T + 1)]
        assert jp[0] == 0 and jp[-1] == n_
        assert all(b - a in (n_ // T, n_ // T + 1) for a, b in zip(jp, jp[1:]))
    acc = np.cumsum(np.r_[0, np.random.default_rng(0).integers(0, 50, 200)]).astype(np.int64)
    out = np.zeros(9, np.int64)
    L.orc_job_pointers_costs(200, oracle.p(acc), 8, oracle.p(out))
    assert out[0] == 0 and out[-1] == 200 and np.all(np.diff(out) >= 0)
    per = np.diff(acc[out])
    assert per.max() - acc[-1] / 8 <= 50  # balanced to within one row's cost


def test_pinned_oracle_runs_keep_the_callers_cpu_set():
    """The CPU baseline pins the oracle's workers (orc_set_pinning); the calling thread
    gets its own CPU set back after every fork-join, so a later measurement still sees
    the process's CPU share, and repeated pinned runs spread over all of it."""
    import os

    before = os.sched_getaffinity(0)
    rng = np.random.default_rng(8)
    m, n, k = 400, 300, 3
    rp, ci, val = random_csr_np(m, n, 0.05, seed=8)
    X = rng.uniform(-1, 1, (n, k))
    lib = oracle.lib(True)
    lib.orc_set_pinning(1)
    try:
        outs = []
        for _ in range(3):
            Y = np.zeros((m, k))
            oracle.spmm(rp, ci, val, m, n, 1.0, X, k, 0.0, Y, k, k, threads=4, fast=True)
            outs.append(Y)
            assert os.sched_getaffinity(0) == before
    finally:
        lib.orc_set_pinning(0)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
