"""Pins the oracle's block-CSR product (KernelMatrixCSR<DenseBlockKernel_Tiny>::Dot,
KernelMatrixCSR.hpp:153-227) and FillLowerTriangleFromUpperTriangle (:71-128).  The reference
has no test of either; these checks are fixed by the mathematics:
  * the block product equals the scalar CSR product (the already pinned oracle spmm) of the
    matrix whose entry (i*R + r, j*C + c) is a_k(r, c) -- bit-exact with integer data;
  * the storage orders (a_RM, x_RM, y_RM) only permute memory: transposing the stored
    blocks / chunks and back gives the same numbers;
  * alpha / beta / empty block rows follow the combine_buffers table;
  * filling the lower triangle makes the dense matrix symmetric when the pattern is."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle


def block_pattern(mb, nb, density, seed, empty_rows=True):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(mb):
        cols = np.flatnonzero(rng.random(nb) < density)
        if empty_rows and i % 7 == 3:
            cols = cols[:0]
        rows.append(cols)
    rp = np.zeros(mb + 1, np.int64)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    return rp, ci


def expand(rp, ci, R, C, A, a_rm=True):
    """Scalar CSR of the block matrix (entry (i*R + r, j*C + c) = a_k(r, c))."""
    mb = len(rp) - 1
    I, J, V = [], [], []
    for i in range(mb):
        for kb in range(rp[i], rp[i + 1]):
            blk = A[kb * R * C:(kb + 1) * R * C]
            blk = blk.reshape(R, C) if a_rm else blk.reshape(C, R).T
            for r in range(R):
                for c in range(C):
                    I.append(i * R + r)
                    J.append(ci[kb] * C + c)
                    V.append(blk[r, c])
    return I, J, np.array(V)


def to_chunks_t(M, nblk, S, k):
    """Row-major (nblk*S) x k -> chunks stored transposed (k x S each)."""
    return np.ascontiguousarray(M.reshape(nblk, S, k).transpose(0, 2, 1)).reshape(-1)


@pytest.mark.parametrize("R,C,k", [(1, 1, 1), (2, 3, 4), (3, 3, 1), (4, 2, 5), (6, 6, 3), (8, 8, 2), (5, 7, 3)])
def test_block_product_equals_expanded_scalar_csr(R, C, k):
    mb, nb = 23, 19
    rp, ci = block_pattern(mb, nb, 0.3, R * 100 + C)
    rng = np.random.default_rng(R + C + k)
    A = rng.integers(-4, 5, int(rp[-1]) * R * C).astype(np.float64)
    X = rng.integers(-4, 5, (nb * C, k)).astype(np.float64)
    Y0 = rng.integers(-4, 5, (mb * R, k)).astype(np.float64)
    for alpha, beta in [(1.0, 0.0), (2.0, 1.0), (-1.0, 3.0)]:
        Y = Y0.copy()
        oracle.bsr_spmm(rp, ci, R, C, A, alpha, X, beta, Y, k)
        I, J, V = expand(rp, ci, R, C, A)
        S = sp.csr_matrix((V, (I, J)), shape=(mb * R, nb * C))
        S.sort_indices()
        Ys = Y0.copy()
        oracle.spmm(S.indptr, S.indices, S.data, mb * R, nb * C, alpha, X, k, beta, Ys, k, k)
        # scalar spmm leaves empty rows to beta as well; integer data: exact in any order
        assert np.array_equal(Y, Ys)


@pytest.mark.parametrize("a_rm,x_rm,y_rm", [(a, x, y) for a in (0, 1) for x in (0, 1) for y in (0, 1)])
def test_storage_orders_only_permute_memory(a_rm, x_rm, y_rm):
    R, C, k, mb, nb = 3, 2, 4, 17, 11
    rp, ci = block_pattern(mb, nb, 0.4, 5)
    rng = np.random.default_rng(9)
    A = rng.integers(-3, 4, int(rp[-1]) * R * C).astype(np.float64)
    X = rng.integers(-3, 4, (nb * C, k)).astype(np.float64)
    Y0 = rng.integers(-3, 4, (mb * R, k)).astype(np.float64)
    Yref = Y0.copy()
    oracle.bsr_spmm(rp, ci, R, C, A, 1.5, X, -0.5, Yref, k)
    A_s = A if a_rm else np.ascontiguousarray(A.reshape(-1, R, C).transpose(0, 2, 1)).reshape(-1)
    X_s = X if x_rm else to_chunks_t(X, nb, C, k)
    Y_s = Y0.copy() if y_rm else to_chunks_t(Y0, mb, R, k)
    oracle.bsr_spmm(rp, ci, R, C, A_s, 1.5, np.ascontiguousarray(X_s), -0.5, Y_s, k, a_rm, x_rm, y_rm)
    Y_back = Y_s.reshape(mb * R, k) if y_rm else Y_s.reshape(mb, k, R).transpose(0, 2, 1).reshape(mb * R, k)
    assert np.array_equal(Y_back, Yref)


def test_random_values_match_dense_product():
    R, C, k, mb, nb = 4, 4, 6, 40, 35
    rp, ci = block_pattern(mb, nb, 0.2, 3)
    rng = np.random.default_rng(4)
    A = rng.normal(size=int(rp[-1]) * R * C)
    X = rng.normal(size=(nb * C, k))
    Y = np.zeros((mb * R, k))
    oracle.bsr_spmm(rp, ci, R, C, A, 1.0, X, 0.0, Y, k)
    I, J, V = expand(rp, ci, R, C, A)
    D = np.zeros((mb * R, nb * C))
    np.add.at(D, (I, J), V)
    assert np.abs(Y - D @ X).max() <= 1e-12 * np.abs(D @ X).max()


def test_alpha_zero_empty_rows_and_beta_flags():
    R, C, k, mb, nb = 2, 2, 3, 9, 9
    rp, ci = block_pattern(mb, nb, 0.5, 8)
    assert np.any(np.diff(rp) == 0)
    A = np.ones(int(rp[-1]) * R * C)
    X = np.ones((nb * C, k))
    Y = np.full((mb * R, k), np.nan)
    oracle.bsr_spmm(rp, ci, R, C, A, 0.0, X, 0.0, Y, k)  # alpha == 0, beta == 0: zeros, NaN dropped
    assert np.array_equal(Y, np.zeros_like(Y))
    Y = np.full((mb * R, k), 2.0)
    oracle.bsr_spmm(rp, ci, R, C, A, 0.0, X, 3.0, Y, k)
    assert np.array_equal(Y, np.full_like(Y, 6.0))
    Y = np.full((mb * R, k), 2.0)
    oracle.bsr_spmm(rp, ci, R, C, A, 1.0, X, 1.0, Y, k)
    empty = np.repeat(np.diff(rp) == 0, R)
    assert np.array_equal(Y[empty], np.full((empty.sum(), k), 2.0))  # beta = 1: untouched
    counts = np.repeat(np.diff(rp), R)[~empty]
    assert np.array_equal(Y[~empty], (2.0 + 2.0 * counts)[:, None] * np.ones(k))


def symmetric_block_pattern(mb, density, seed, drop_mirror=None):
    rng = np.random.default_rng(seed)
    M = rng.random((mb, mb)) < density
    M = M | M.T | np.eye(mb, dtype=bool)
    if drop_mirror is not None:
        i, j = drop_mirror  # keep (i, j) below the diagonal, remove its mirror (j, i)
        M[i, j] = True
        M[j, i] = False
    rp = np.zeros(mb + 1, np.int64)
    rp[1:] = np.cumsum(M.sum(1))
    ci = np.concatenate([np.flatnonzero(M[i]) for i in range(mb)]).astype(np.int32)
    return rp, ci


def test_fill_lower_makes_the_matrix_symmetric():
    R, mb = 3, 30
    rp, ci = symmetric_block_pattern(mb, 0.15, 2)
    A = np.random.default_rng(6).normal(size=int(rp[-1]) * R * R)
    rows = np.repeat(np.arange(mb), np.diff(rp))
    for kb in np.flatnonzero(ci == rows):  # diagonal blocks symmetric (the fill leaves them)
        b = A[kb * R * R:(kb + 1) * R * R].reshape(R, R)
        A[kb * R * R:(kb + 1) * R * R] = ((b + b.T) / 2).reshape(-1)
    upper = A.copy()
    oracle.bsr_fill_lower(rp, ci, R, A)
    I, J, V = expand(rp, ci, R, R, A)
    D = np.zeros((mb * R, mb * R))
    D[I, J] = V
    assert np.array_equal(D, D.T)
    # blocks on and above the diagonal are untouched
    keep = np.repeat(ci >= rows, R * R)
    assert np.array_equal(A[keep], upper[keep])


def test_fill_lower_leaves_blocks_without_a_mirror():
    R, mb = 2, 12
    rp, ci = symmetric_block_pattern(mb, 0.2, 3, drop_mirror=(9, 4))
    A = np.arange(int(rp[-1]) * R * R, dtype=np.float64)
    before = A.copy()
    oracle.bsr_fill_lower(rp, ci, R, A)
    kb = rp[9] + np.flatnonzero(ci[rp[9]:rp[10]] == 4)[0]
    assert np.array_equal(A[kb * 4:(kb + 1) * 4], before[kb * 4:(kb + 1) * 4])
