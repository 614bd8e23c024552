"""GPU parity of tg_csr_spmm (the HIP SpMV/SpMM behind MatrixCSR::Dot) against the oracle.

Bars: FP64 within 1e-12 normwise relative error (max|Y-Y_ref|/max|Y_ref|, the reference's
metric) and inside the per-entry rounding bound gamma_{L+3}(|alpha||A||X| + |beta||Y|);
FP32 inside the same bound with u = 2^-24 and within sqrt(eps_f32) normwise (the
reference's own rule, Unit_Test_SpMM/main.cpp:144).  Structure-free cases with small
integers are compared bit for bit.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle
from helpers import dense_of, random_csr_np, rel_err, spmm_bound
from tensors_amd import _capi as capi
from tensors_amd import generators as gen
from tensors_amd.sparse import MatrixCSR, spmm_raw

pytestmark = pytest.mark.gpu

ALPHAS = [-1.0, 0.0, 1.0, 2.0, np.pi]
BETAS = [-1.0, 0.0, 1.0, np.pi]


def to_dev(a, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return t if dtype is None else t.to(dtype)


def run_gpu(dev, rp, ci, val, m, n, alpha, X, beta, Y0, k, rp_dtype=torch.int64, NRHS=0, use_jp=False,
            ldX=None, ldY=None):
    dt = torch.float64 if X.dtype == np.float64 else torch.float32
    rpd = to_dev(rp, dev).to(rp_dtype)
    cid = to_dev(ci, dev)
    vd = None if val is None else to_dev(val, dev, dt)
    Xd = to_dev(X, dev, dt)
    Yd = to_dev(Y0, dev, dt)
    jp, blocks = None, 0
    if use_jp:
        blocks = capi.lib().tg_csr_default_blocks(m, int(rp[m]), k)
        jp = torch.empty(blocks + 1, dtype=torch.int32, device=dev)
        capi.call("tg_csr_job_pointers", m, rpd.data_ptr(), capi.I64 if rp_dtype == torch.int64 else capi.I32,
                  blocks, jp.data_ptr(), 0)
    st = spmm_raw(m, n, int(rp[m]), rpd, cid, vd, alpha, Xd, ldX or X.shape[1], beta, Yd, ldY or Y0.shape[1], k,
                  NRHS, jp, blocks, None)
    torch.cuda.synchronize()
    return st, Yd.cpu().numpy()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8, 16, 32, 64, 100, 130])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmm_alpha_beta_grid(cuda, k, dtype):
    m, n = 300, 257
    rp, ci, val = random_csr_np(m, n, 0.06, seed=k, empty_rows=0.1, dtype=dtype)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, (n, k)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (m, k)).astype(dtype)
    u = 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24
    for alpha, beta in itertools.product(ALPHAS, BETAS):
        st, Y = run_gpu(cuda, rp, ci, val, m, n, alpha, X, beta, Y0, k)
        assert st == 0
        Yo = Y0.copy()
        assert oracle.spmm(rp, ci, val, m, n, alpha, X, k, beta, Yo, k, k) == 0
        bound = spmm_bound(m, n, rp, ci, val.astype(np.float64), alpha, X.astype(np.float64), beta,
                           Y0.astype(np.float64), u)
        exact = alpha * (dense_of(m, n, rp, ci, val.astype(np.float64)) @ X.astype(np.float64)) \
            + (beta * Y0.astype(np.float64) if beta != 0 else 0)
        if alpha == 0.0:
            exact = beta * Y0.astype(np.float64)
        assert np.all(np.abs(Y.astype(np.float64) - exact) <= bound), (alpha, beta)
        tol = 1e-12 if dtype == np.float64 else np.sqrt(np.finfo(np.float32).eps)
        assert rel_err(Y, Yo) <= tol, (alpha, beta, rel_err(Y, Yo))


@pytest.mark.parametrize("rp_dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("use_jp", [False, True])
@pytest.mark.parametrize("k", [1, 4, 16])
def test_index_types_and_partition(cuda, rp_dtype, use_jp, k):
    m, n = 2000, 1500
    rp, ci, val = random_csr_np(m, n, 0.01, seed=3, empty_rows=0.05)
    rng = np.random.default_rng(0)
    X = rng.integers(-6, 7, (n, k)).astype(np.float64)
    Y0 = rng.integers(-6, 7, (m, k)).astype(np.float64)
    vi = rng.integers(-3, 4, len(ci)).astype(np.float64)
    st, Y = run_gpu(cuda, rp, ci, vi, m, n, 2.0, X, -1.0, Y0, k, rp_dtype=rp_dtype, use_jp=use_jp)
    assert st == 0
    Yo = Y0.copy()
    oracle.spmm(rp, ci, vi, m, n, 2.0, X, k, -1.0, Yo, k, k)
    assert np.array_equal(Y, Yo)  # exact integer arithmetic: bit-exact


def test_binary_matrix(cuda):
    m, n, k = 500, 400, 8
    rp, ci, _ = random_csr_np(m, n, 0.02, seed=4, empty_rows=0.1)
    X = np.random.default_rng(1).integers(-9, 10, (n, k)).astype(np.float64)
    Y0 = np.full((m, k), np.nan)
    st, Y = run_gpu(cuda, rp, ci, None, m, n, 1.0, X, 0.0, Y0, k)
    assert st == 0
    assert np.array_equal(Y, dense_of(m, n, rp, ci, None) @ X)


def test_beta_zero_nan_and_empty_rows(cuda):
    m, n = 400, 300
    rp, ci, val = random_csr_np(m, n, 0.02, seed=5, empty_rows=0.3)
    empty = np.diff(rp) == 0
    for k in (1, 4, 16):
        X = np.ones((n, k))
        st, Y = run_gpu(cuda, rp, ci, val, m, n, 1.0, X, 0.0, np.full((m, k), np.nan), k)
        assert st == 0 and np.all(np.isfinite(Y)) and np.all(Y[empty] == 0.0)
        st, Y = run_gpu(cuda, rp, ci, val, m, n, 1.0, X, 1.0, np.full((m, k), np.inf), k)
        assert np.all(np.isinf(Y[empty]))


def test_strided_and_misaligned(cuda):
    m, n, k, ld = 333, 222, 6, 11
    rp, ci, val = random_csr_np(m, n, 0.05, seed=6)
    rng = np.random.default_rng(2)
    Xf = rng.uniform(-1, 1, (n, ld))
    Yf = rng.uniform(-1, 1, (m, ld))
    st, Y = run_gpu(cuda, rp, ci, val, m, n, 1.5, Xf, -0.5, Yf, k, ldX=ld, ldY=ld)
    assert st == 0
    Yo = Yf.copy()
    oracle.spmm(rp, ci, val, m, n, 1.5, Xf, ld, -0.5, Yo, ld, k)
    assert rel_err(Y[:, :k], Yo[:, :k]) <= 1e-12
    assert np.array_equal(Y[:, k:], Yf[:, k:])
    # X starting one element off a 16-byte boundary forces the scalar-load path
    dev = cuda
    Xbuf = torch.zeros(n * 4 + 1, dtype=torch.float64, device=dev)
    Xv = Xbuf[1:].view(n, 4)
    Xv.copy_(torch.from_numpy(Xf[:, :4]))
    Yd = torch.zeros(m, 4, dtype=torch.float64, device=dev)
    st = spmm_raw(m, n, int(rp[m]), to_dev(rp, dev), to_dev(ci, dev), to_dev(val, dev), 1.0, Xv, 4, 0.0, Yd,
                  4, 4)
    Yo = np.zeros((m, 4))
    oracle.spmm(rp, ci, val, m, n, 1.0, np.ascontiguousarray(Xf[:, :4]), 4, 0.0, Yo, 4, 4)
    assert st == 0 and rel_err(Yd.cpu().numpy(), Yo) <= 1e-12


def test_noop_error_cases(cuda):
    m = n = 64
    rp, ci, val = random_csr_np(m, n, 0.1, seed=7)
    dev = cuda
    Xd = torch.ones(n, 2, dtype=torch.float64, device=dev)
    snap = Xd.clone()
    st = spmm_raw(m, n, int(rp[m]), to_dev(rp, dev), to_dev(ci, dev), to_dev(val, dev), 1.0, Xd, 2, 0.0, Xd, 2,
                  2)
    assert st == capi.ERR_ALIAS and torch.equal(Xd, snap)
    Yd = torch.full((m, 2), 7.0, dtype=torch.float64, device=dev)
    st = spmm_raw(m, n, int(rp[m]), to_dev(rp, dev), to_dev(ci, dev), to_dev(val, dev), 1.0, Xd, 2, 0.0, Yd, 2,
                  2, NRHS=4)
    assert st == capi.ERR_NRHS and torch.all(Yd == 7.0)
    # empty matrix (nnz == 0): every row empty
    rp0 = np.zeros(m + 1, np.int64)
    st, Y = run_gpu(dev, rp0, np.zeros(0, np.int32), np.zeros(0), m, n, 1.0, np.ones((n, 3)), 0.0,
                    np.full((m, 3), np.nan), 3)
    assert st == 0 and np.all(Y == 0.0)
    # m == 0 and nrhs == 0 are no-ops
    assert capi.lib().tg_csr_spmm(0, n, 0, to_dev(rp0, dev).data_ptr(), capi.I64, None, None, capi.F64, 1.0,
                                  Xd.data_ptr(), 2, 0.0, Yd.data_ptr(), 2, 2, 0, None, 0, None) == 0


def test_grid_laplacian_256_config(cuda):
    """GridLaplacian(256), FP64, nrhs 1 and 16, alpha = 1, beta = 0 (configs[0] shape)."""
    rp, ci, val = oracle.grid_laplacian(256, 1.0)
    m = n = 256 * 256
    assert len(ci) == 326656
    for k in (1, 16):
        X = gen.uniform(1001, n * k).reshape(n, k)
        st, Y = run_gpu(cuda, rp, ci, val, m, n, 1.0, X, 0.0, np.zeros((m, k)), k)
        Yo = np.zeros((m, k))
        oracle.spmm(rp, ci, val, m, n, 1.0, X, k, 0.0, Yo, k, k)
        assert st == 0 and rel_err(Y, Yo) <= 1e-12


@pytest.mark.parametrize("k", [1, 4, 16, 64])
def test_random_csr_1e5(cuda, k):
    """configs[1] generator at n = 1e5 (oracle-sized): FP64 and FP32."""
    m = n = 100_000
    rp, ci, v64, v32 = gen.random_csr(m, n, seed=2002, want32=True)
    X = gen.uniform(2002, n * k).reshape(n, k)
    Yo = np.zeros((m, k))
    oracle.spmm(rp, ci, v64, m, n, 1.0, X, k, 0.0, Yo, k, k, threads=8)
    st, Y = run_gpu(cuda, rp, ci, v64, m, n, 1.0, X, 0.0, np.zeros((m, k)), k)
    assert st == 0 and rel_err(Y, Yo) <= 1e-12
    X32 = X.astype(np.float32)
    st, Y32 = run_gpu(cuda, rp, ci, v32, m, n, 1.0, X32, 0.0, np.zeros((m, k), np.float32), k)
    Yo32 = np.zeros((m, k), np.float32)
    oracle.spmm(rp, ci, v32, m, n, 1.0, X32, k, 0.0, Yo32, k, k, threads=8)
    assert st == 0 and rel_err(Y32, Yo32) <= 1e-5


@pytest.mark.parametrize("k", [1, 16, 64])
def test_random_csr_full_size_properties(cuda, k):
    """n = 1e6 (full configs[1] size): linearity A(X1 + X2) = A X1 + A X2 and a sampled
    set of rows checked against the oracle on the corresponding sub-matrix."""
    m = n = 1_000_000
    rp, ci, v64, _ = gen.random_csr(m, n, seed=2002)
    dev = cuda
    rpd, cid, vd = to_dev(rp, dev), to_dev(ci, dev), to_dev(v64, dev)
    X1 = torch.from_numpy(gen.uniform(7, n * k).reshape(n, k)).to(dev)
    X2 = torch.from_numpy(gen.uniform(8, n * k).reshape(n, k)).to(dev)
    Y1 = torch.empty(m, k, dtype=torch.float64, device=dev)
    Y2 = torch.empty_like(Y1)
    Y3 = torch.empty_like(Y1)
    nnz = int(rp[m])
    for X, Y in ((X1, Y1), (X2, Y2), (X1 + X2, Y3)):
        assert spmm_raw(m, n, nnz, rpd, cid, vd, 1.0, X, k, 0.0, Y, k, k) == 0
    torch.cuda.synchronize()
    lin = (Y3 - Y1 - Y2).abs().max().item() / Y3.abs().max().item()
    assert lin < 1e-13
    rows = np.sort(np.random.default_rng(0).choice(m, 2000, replace=False))
    srp = np.r_[0, np.cumsum(rp[rows + 1] - rp[rows])].astype(np.int64)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
    Yo = np.zeros((len(rows), k))
    X1h = X1.cpu().numpy()
    oracle.spmm(srp, ci[idx], v64[idx], len(rows), n, 1.0, X1h, k, 0.0, Yo, k, k)
    assert rel_err(Y1.cpu().numpy()[rows], Yo) <= 1e-12


def test_matrixcsr_dot_overloads(cuda):
    """Python mirror of MatrixCSR::Dot: Tensor1 / Tensor2 overloads and the shape check."""
    m, n = 120, 90
    rp, ci, val = random_csr_np(m, n, 0.1, seed=10)
    A = MatrixCSR(rp, ci, val, m, n)
    x = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, n)).to(cuda)
    y = torch.zeros(m, dtype=torch.float64, device=cuda)
    assert A.Dot(1.0, x, 0.0, y) == 0
    assert rel_err(y.cpu().numpy(), dense_of(m, n, rp, ci, val) @ x.cpu().numpy()) <= 1e-12
    X = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, (n, 5))).to(cuda)
    Y = torch.zeros(m, 5, dtype=torch.float64, device=cuda)
    assert A.Dot(2.0, X, 0.0, Y, use_job_ptr=True) == 0
    assert rel_err(Y.cpu().numpy(), 2.0 * dense_of(m, n, rp, ci, val) @ X.cpu().numpy()) <= 1e-12
    Ybad = torch.full((m + 1, 5), 3.0, dtype=torch.float64, device=cuda)
    assert A.Dot(1.0, X, 0.0, Ybad) == capi.ERR_SHAPE and torch.all(Ybad == 3.0)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_kernel_variants(cuda, variant, dtype):
    """Both nrhs == 1 kernels (streaming and team-per-row) against the oracle, including
    rows longer than one streaming chunk and blocks of more than 1024 rows."""
    L = capi.lib()
    old = L.tg_tune_get(b"spmv_kernel")
    try:
        assert L.tg_tune_set(b"spmv_kernel", variant) == 0
        # long rows (up to 5000 nonzeros) next to many empty / short rows
        rng = np.random.default_rng(11)
        m, n = 6000, 7000
        lens = rng.integers(0, 4, m)
        lens[rng.choice(m, 5, replace=False)] = rng.integers(2500, 5000, 5)
        rp = np.r_[0, np.cumsum(lens)].astype(np.int64)
        ci = np.concatenate([np.sort(rng.choice(n, L_, replace=False)) for L_ in lens]).astype(np.int32)
        val = rng.normal(size=len(ci)).astype(dtype)
        X = rng.uniform(-1, 1, (n, 1)).astype(dtype)
        Y0 = rng.uniform(-1, 1, (m, 1)).astype(dtype)
        for alpha, beta in ((1.0, 0.0), (-2.0, 1.0), (0.5, np.pi)):
            st, Y = run_gpu(cuda, rp, ci, val, m, n, alpha, X, beta, Y0, 1)
            Yo = Y0.copy()
            oracle.spmm(rp, ci, val, m, n, alpha, X, 1, beta, Yo, 1, 1)
            tol = 1e-12 if dtype == np.float64 else 1e-5
            assert st == 0 and rel_err(Y, Yo) <= tol
    finally:
        L.tg_tune_set(b"spmv_kernel", old)
