"""GPU parity of the block-CSR product (tg_bsr_spmm, KernelMatrixCSR<DenseBlockKernel_Tiny>::Dot,
KernelMatrixCSR.hpp:153-227) and of FillLowerTriangleFromUpperTriangle (:71-128) against the
oracle.  Integer-valued data make every product exact, so those cases compare bit for bit;
random values are compared within the gamma bound of a block row's sum."""
import numpy as np
import pytest
import torch

import oracle
from test_oracle_bsr import block_pattern, symmetric_block_pattern
from tensors_amd.sparse import KernelMatrixCSR

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["vector", "scalar"])
def mapping(request):
    """Both lane mappings of tg_bsr_spmm (knob bsr_vector), restored afterwards."""
    from tensors_amd import _capi as capi
    before = int(capi.lib().tg_tune_get(b"bsr_vector"))
    capi.call("tg_tune_set", b"bsr_vector", request.param)
    yield request.param
    capi.call("tg_tune_set", b"bsr_vector", before)


def dev(a, cuda, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dtype)


@pytest.mark.parametrize("R,C", [(1, 1), (2, 2), (3, 3), (4, 4), (6, 6), (8, 8), (2, 5), (5, 3), (12, 12)])
@pytest.mark.parametrize("k", [1, 3, 8, 20])
def test_bsr_integer_data_bit_exact(cuda, mapping, R, C, k):
    mb, nb = 57, 49
    rp, ci = block_pattern(mb, nb, 0.15, R * 31 + C)
    rng = np.random.default_rng(R * C + k)
    A = rng.integers(-4, 5, int(rp[-1]) * R * C).astype(np.float64)
    X = rng.integers(-4, 5, (nb * C, k)).astype(np.float64)
    Y0 = rng.integers(-4, 5, (mb * R, k)).astype(np.float64)
    K = KernelMatrixCSR(rp if k % 2 else rp.astype(np.int32), ci, mb, nb, R, C)
    for alpha, beta in [(1.0, 0.0), (2.0, 1.0), (-1.0, 0.5)]:
        Yo = Y0.copy()
        oracle.bsr_spmm(rp, ci, R, C, A, alpha, X, beta, Yo, k)
        for dtype in (torch.float64, torch.float32):
            Yd = dev(Y0, cuda, dtype)
            assert K.Dot(dev(A, cuda, dtype), alpha, dev(X, cuda, dtype), beta, Yd, k) == 0
            assert np.array_equal(Yd.double().cpu().numpy(), Yo), (alpha, beta, dtype)


@pytest.mark.parametrize("a_rm,x_rm,y_rm", [(a, x, y) for a in (0, 1) for x in (0, 1) for y in (0, 1)])
def test_bsr_storage_orders(cuda, mapping, a_rm, x_rm, y_rm):
    R, C, k, mb, nb = 3, 4, 5, 64, 40
    rp, ci = block_pattern(mb, nb, 0.2, 17)
    rng = np.random.default_rng(2)
    A = rng.normal(size=int(rp[-1]) * R * C)
    X = rng.normal(size=nb * C * k)
    Y0 = rng.normal(size=mb * R * k)
    Yo = Y0.copy()
    oracle.bsr_spmm(rp, ci, R, C, A, 0.75, X, -2.0, Yo, k, a_rm, x_rm, y_rm)
    K = KernelMatrixCSR(rp, ci, mb, nb, R, C, a_rm, x_rm, y_rm)
    Yd = dev(Y0, cuda, torch.float64)
    assert K.Dot(dev(A, cuda, torch.float64), 0.75, dev(X, cuda, torch.float64), -2.0, Yd, k) == 0
    Yg = Yd.cpu().numpy()
    # per entry: |error| <= gamma_(C * blocks) * sum |a| |x| + |beta y|, bounded crudely here
    assert np.abs(Yg - Yo).max() <= 1e-13 * max(1.0, np.abs(Yo).max())


def test_bsr_special_cases(cuda, mapping):
    R, C, k, mb, nb = 2, 3, 4, 30, 20
    rp, ci = block_pattern(mb, nb, 0.3, 4)
    assert np.any(np.diff(rp) == 0)
    K = KernelMatrixCSR(rp, ci, mb, nb, R, C)
    A = torch.ones(int(rp[-1]) * R * C, dtype=torch.float64, device=cuda)
    X = torch.ones(nb * C * k, dtype=torch.float64, device=cuda)
    Y = torch.full((mb * R * k,), float("nan"), dtype=torch.float64, device=cuda)
    assert K.Dot(A, 0.0, X, 0.0, Y, k) == 0  # alpha == 0, beta == 0: zeros even over NaN
    assert torch.all(Y == 0)
    Y.fill_(2.0)
    assert K.Dot(A, 0.0, X, 3.0, Y, k) == 0
    assert torch.all(Y == 6.0)
    Y.fill_(float("nan"))
    assert K.Dot(A, 1.0, X, 0.0, Y, k) == 0  # beta == 0 never reads Y; empty rows -> 0
    Yo = np.zeros(mb * R * k)
    oracle.bsr_spmm(rp, ci, R, C, np.ones(int(rp[-1]) * R * C), 1.0, np.ones(nb * C * k), 0.0, Yo, k)
    assert np.array_equal(Y.cpu().numpy(), Yo)


def test_bsr_alias_is_a_no_op(cuda):
    R = C = 2
    mb = nb = 10
    rp, ci = block_pattern(mb, nb, 0.4, 6)
    K = KernelMatrixCSR(rp, ci, mb, nb, R, C)
    A = torch.ones(int(rp[-1]) * 4, dtype=torch.float64, device=cuda)
    Y = torch.arange(mb * R * 3, dtype=torch.float64, device=cuda)
    before = Y.clone()
    assert K.Dot(A, 1.0, Y, 0.0, Y, 3) != 0
    assert torch.equal(Y, before)


def test_bsr_large_against_expanded_csr_spmm(cuda, mapping):
    """A 3 x 3 block 7-point grid pattern (40^3 block rows, 120 x 40^3 scalar unknowns):
    the block product equals the scalar CSR SpMM of the expanded matrix (tg_csr_spmm)."""
    import scipy.sparse as sp

    from tensors_amd import generators as gen
    from tensors_amd.sparse import MatrixCSR

    g, R, k = 40, 3, 4
    rp, ci, _ = gen.laplacian_3d(g, 1.0)
    mb = g ** 3
    rng = np.random.default_rng(8)
    A = rng.normal(size=int(rp[-1]) * R * R)
    X = rng.normal(size=(mb * R, k))
    K = KernelMatrixCSR(rp, ci, mb, mb, R, R)
    Yd = torch.zeros(mb * R, k, dtype=torch.float64, device=cuda)
    assert K.Dot(dev(A, cuda, torch.float64), 1.0, dev(X, cuda, torch.float64), 0.0, Yd, k) == 0
    rows = np.repeat(np.arange(mb), np.diff(rp))
    blk = A.reshape(-1, R, R)
    I = (rows[:, None, None] * R + np.arange(R)[None, :, None]).repeat(R, 2).reshape(-1)
    J = (ci[:, None, None].astype(np.int64) * R + np.arange(R)[None, None, :]).repeat(R, 1).reshape(-1)
    S = sp.csr_matrix((blk.reshape(-1), (I, J)), shape=(mb * R, mb * R))
    S.sort_indices()
    M = MatrixCSR(S.indptr.astype(np.int64), S.indices, S.data, mb * R, mb * R)
    Ys = torch.zeros(mb * R, k, dtype=torch.float64, device=cuda)
    assert M.Dot(1.0, dev(X, cuda, torch.float64), 0.0, Ys) == 0
    ref = S @ X
    assert np.abs(Yd.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(Yd.cpu().numpy() - Ys.cpu().numpy()).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("R", [1, 2, 3, 5])
def test_fill_lower_matches_oracle(cuda, R):
    mb = 80
    rp, ci = symmetric_block_pattern(mb, 0.08, R, drop_mirror=(50, 7))
    A = np.random.default_rng(R).normal(size=int(rp[-1]) * R * R)
    Ao = A.copy()
    oracle.bsr_fill_lower(rp, ci, R, Ao)
    K = KernelMatrixCSR(rp, ci, mb, mb, R, R)
    for dtype in (torch.float64, torch.float32):
        Ad = dev(A, cuda, dtype)
        K.FillLowerTriangleFromUpperTriangle(Ad)
        exp = Ao if dtype == torch.float64 else Ao.astype(np.float32)
        assert np.array_equal(Ad.cpu().numpy(), exp)
