"""Long rows split off the row kernels (tg_csr_spmm_plan with a plan that found rows of
more than spmm_long_row nonzeros): every result against the oracle's SpMM at the bars of
the other SpMM tests (FP64 1e-12 normwise and the per-entry rounding bound, FP32 1e-5), for
the alpha/beta rules (beta = 0 with NaN in Y, beta = 1, general beta), empty rows, nrhs from
1 to 130 (column slabs of the row kernels included), binary matrices and both row-pointer
types; integer data bit for bit; and the same products with the split switched off agree."""
import numpy as np
import pytest
import torch

import oracle
from helpers import rel_err, spmm_bound
from tensors_amd import _capi as capi
from tensors_amd.sparse import SlicePlan

pytestmark = pytest.mark.gpu


def skewed(m, n, seed, dtype=np.float64, integer=False):
    """Short rows (Poisson(20)), some empty, and four long rows of 9 000 - 30 000 nonzeros
    (one to eight 4096-nonzero segments), sorted distinct columns per row."""
    rng = np.random.default_rng(seed)
    lens = rng.poisson(20, m)
    lens[rng.choice(m, m // 20, replace=False)] = 0
    long_rows = rng.choice(m, 4, replace=False)
    lens[long_rows] = [9000, 12288, 20000, 30000]
    rp = np.r_[0, np.cumsum(lens)].astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L_, replace=False)) for L_ in lens]).astype(np.int32)
    if integer:
        val = rng.integers(-3, 4, len(ci)).astype(dtype)
    else:
        val = rng.normal(size=len(ci)).astype(dtype)
    return rp, ci, val, np.sort(long_rows)


@pytest.fixture
def long_knob():
    before = int(capi.lib().tg_tune_get(b"spmm_long_row"))
    yield lambda v: capi.call("tg_tune_set", b"spmm_long_row", int(v))
    capi.call("tg_tune_set", b"spmm_long_row", before)


def plan_product(cuda, rp, ci, val, m, n, alpha, X, beta, Y0, k, rp_dtype=torch.int32, binary=False):
    dt = torch.float64 if X.dtype == np.float64 else torch.float32
    rpd = torch.from_numpy(rp).to(cuda).to(rp_dtype)
    cid = torch.from_numpy(ci).to(cuda)
    plan = SlicePlan(m, n, int(rp[-1]), rpd, cid)
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(cuda)
    Yd = torch.from_numpy(np.ascontiguousarray(Y0)).to(cuda)
    vd = None if binary else torch.from_numpy(val).to(cuda).to(dt)
    st = plan.spmm(vd, alpha, Xd, k, beta, Yd, k, k)
    torch.cuda.synchronize()
    nl = plan.long_rows()
    del plan
    return st, Yd.cpu().numpy(), nl


@pytest.mark.parametrize("k", [1, 2, 5, 16, 64, 130])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_rows_match_oracle(cuda, k, dtype):
    m, n = 3000, 40000
    rp, ci, val, _ = skewed(m, n, seed=100 + k, dtype=dtype)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, (n, k)).astype(dtype)
    u = 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24
    for alpha, beta in ((1.0, 0.0), (-2.0, 1.0), (np.pi, -1.0), (0.5, np.pi)):
        Y0 = rng.uniform(-1, 1, (m, k)).astype(dtype)
        if beta == 0.0:
            Y0[::7] = np.nan  # beta = 0 never reads Y
        st, Y, nl = plan_product(cuda, rp, ci, val, m, n, alpha, X, beta, Y0, k)
        print(f"long rows: k {k} {dtype.__name__} alpha {alpha} beta {beta}: {nl} long rows")
        assert st == 0 and nl == 4
        Yo = Y0.copy()
        oracle.spmm(rp, ci, val, m, n, alpha, X, k, beta, Yo, k, k)
        assert not np.isnan(Y).any()
        if dtype == np.float64:
            assert rel_err(Y, Yo) <= 1e-12
        else:
            assert rel_err(Y, Yo) <= 1e-5
        bound = spmm_bound(m, n, rp, ci, val, alpha, X, beta, np.nan_to_num(Y0), u)
        assert np.all(np.abs(Y - Yo) <= 2 * bound + 1e-300)


@pytest.mark.parametrize("rp_dtype", [torch.int32, torch.int64], ids=["rp32", "rp64"])
@pytest.mark.parametrize("binary", [False, True], ids=["valued", "binary"])
def test_long_rows_integer_data_bit_exact(cuda, rp_dtype, binary):
    """Small integers: every sum is exact in FP64 in any order, so the split product equals
    the oracle's bit for bit, for both row-pointer types and a binary matrix (a = 1)."""
    m, n, k = 2000, 35000, 3
    rp, ci, val, _ = skewed(m, n, seed=7, integer=True)
    rng = np.random.default_rng(8)
    X = rng.integers(-4, 5, (n, k)).astype(np.float64)
    Y0 = rng.integers(-4, 5, (m, k)).astype(np.float64)
    st, Y, nl = plan_product(cuda, rp, ci, val, m, n, 2.0, X, 1.0, Y0, k, rp_dtype=rp_dtype, binary=binary)
    assert st == 0 and nl == 4
    Yo = Y0.copy()
    oracle.spmm(rp, ci, np.ones(len(ci)) if binary else val, m, n, 2.0, X, k, 1.0, Yo, k, k)
    assert np.array_equal(Y, Yo)


def test_long_rows_switched_off_agree(cuda, long_knob):
    """spmm_long_row = 0 keeps every row in the row kernels: the same product within
    rounding, and no long rows in the plan."""
    m, n, k = 3000, 40000, 4
    rp, ci, val, _ = skewed(m, n, seed=9)
    X = np.random.default_rng(10).uniform(-1, 1, (n, k))
    Y0 = np.zeros((m, k))
    st1, Y1, nl1 = plan_product(cuda, rp, ci, val, m, n, 1.0, X, 0.0, Y0, k)
    long_knob(0)
    st0, Y0b, nl0 = plan_product(cuda, rp, ci, val, m, n, 1.0, X, 0.0, Y0, k)
    assert st1 == st0 == 0 and nl1 == 4 and nl0 == 0
    assert rel_err(Y1, Y0b) <= 1e-13
