"""CPU pins of the oracle's structure routines against scipy.sparse, which the GPU
structure kernels are then compared with bit for bit (tests/test_gpu_structure.py):
transpose equals scipy's A.T (sorted columns per row, values carried along), the value
product equals scipy's A @ B and the pattern product scipy's boolean product."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle


def random_csr(m, n, count, seed):
    rng = np.random.default_rng(seed)
    I = rng.integers(0, m, count).astype(np.int32)
    J = rng.integers(0, n, count).astype(np.int32)
    V = rng.standard_normal(count)
    return oracle.from_triples(I, J, V, m, n)


@pytest.mark.parametrize("m,n,count", [(230, 170, 3000), (1, 40, 30), (50, 1, 20), (300, 300, 0)])
def test_transpose_equals_scipy(m, n, count):
    rp, ci, val = random_csr(m, n, count, seed=m + n)
    trp, tci, tval = oracle.transpose(m, n, rp, ci, val)
    T = sp.csr_matrix((val, ci, rp), shape=(m, n)).T.tocsr()
    T.sort_indices()
    assert np.array_equal(trp, T.indptr) and np.array_equal(tci, T.indices) and np.array_equal(tval, T.data)
    # pattern-only transpose gives the same structure
    prp, pci, _ = oracle.transpose(m, n, rp, ci, None)
    assert np.array_equal(prp, trp) and np.array_equal(pci, tci)


@pytest.mark.parametrize("values", [True, False])
def test_spgemm_equals_scipy(values):
    m, k, n = 120, 90, 110
    A = random_csr(m, k, 900, seed=4)
    B = random_csr(k, n, 800, seed=5)
    crp, cci, cval = oracle.spgemm(m, k, n, A, B, values=values)
    As = sp.csr_matrix((A[2], A[1], A[0]), shape=(m, k))
    Bs = sp.csr_matrix((B[2], B[1], B[0]), shape=(k, n))
    if values:
        C = (As @ Bs).tocsr()
        C.sort_indices()
        got = sp.csr_matrix((cval, cci, crp), shape=(m, n))
        assert np.array_equal(crp, C.indptr) and np.array_equal(cci, C.indices)
        assert np.abs(got.toarray() - C.toarray()).max() <= 1e-13 * np.abs(C.toarray()).max()
    else:
        P = (As.astype(bool) @ Bs.astype(bool)).tocsr()
        P.sort_indices()
        assert np.array_equal(crp, P.indptr) and np.array_equal(cci, P.indices)
