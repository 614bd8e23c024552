"""Batched CG engine: every system's solution matches the FP64 oracle CG, mixed systems
reach the true 1e-10 residual, and a system's result is bitwise the same alone, in a mix
and when admitted mid-flight (continuous batching)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.batch import BatchCG

pytestmark = pytest.mark.gpu


def systems(count, seed=4004):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        n_target = int(np.exp(rng.uniform(np.log(300), np.log(6000))))
        dims = 2 if rng.random() < 0.5 else 3
        side = max(2, int(round(n_target ** (1.0 / dims))))
        sides = (side,) * dims
        rp, ci, val = gen.weighted_grid(dims, sides, 1e-3, seed=seed + s)
        n = len(rp) - 1
        k = int(rng.integers(1, 9))
        B = gen.uniform(seed * 10 + s, n * k).reshape(n, k)
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
        out.append((rp, ci, val, d, B, n, k))
    return out


def to_dev(sysd, dev):
    rp, ci, val, d, B, n, k = sysd
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    return (T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64), T(d, torch.float64),
            T(B, torch.float64), torch.zeros(n, k, dtype=torch.float64, device=dev))


def run_batch(sysl, dev, mixed, tol, order=None, staged=False):
    bt = BatchCG()
    devs = [to_dev(s, dev) for s in sysl]
    ids = []
    idx = order if order is not None else range(len(sysl))
    half = len(sysl) // 2
    for j, i in enumerate(idx):
        rp, ci, val, d, B, X = devs[i]
        ids.append((i, bt.add(rp, ci, val, B, X, tol, 20000, diag=d, mixed=mixed)))
        if staged and j == half:
            bt.step(3)  # admit the rest while the first half is running
    bt.run(8)
    res = {i: bt.result(sid) for i, sid in ids}
    X = {i: devs[i][5].cpu().numpy() for i, _ in ids}
    bt.close()
    return X, res


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_matches_oracle(cuda, mixed):
    sysl = systems(10)
    X, res = run_batch(sysl, cuda, mixed, 1e-10)
    for i, (rp, ci, val, d, B, n, k) in enumerate(sysl):
        assert res[i]["succeeded"], (i, res[i])
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        true = np.linalg.norm(B - A @ X[i], axis=0) / np.linalg.norm(B, axis=0)
        if mixed:
            # mixed systems stop on the FP64 true residual
            assert np.all(true <= 1.5e-10), (i, true)
            Xs = np.linalg.solve(A.toarray(), B) if n <= 3000 else None
            if Xs is not None:
                assert rel_err(X[i], Xs) < 1e-6
        else:
            # plain FP64 systems run the reference recurrences: same iteration count as the
            # oracle CG at the same tolerance, and the same iterate up to rounding
            Xo = np.zeros((n, k))
            ok, it, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 20000, 1e-10, prec_diag=d, threads=4)
            assert ok
            assert abs(res[i]["iterations"] - it) <= 1, (i, res[i]["iterations"], it)
            assert rel_err(X[i], Xo) < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_alone_equals_mix_and_midflight(cuda, mixed):
    sysl = systems(8, seed=777)
    Xmix, _ = run_batch(sysl, cuda, mixed, 1e-10)
    Xstaged, _ = run_batch(sysl, cuda, mixed, 1e-10, order=list(reversed(range(8))), staged=True)
    for j in (0, 3, 7):
        Xalone, _ = run_batch([sysl[j]], cuda, mixed, 1e-10)
        assert np.array_equal(Xalone[0], Xmix[j])
        assert np.array_equal(Xstaged[j], Xmix[j])


def test_batch_spmm_kernels_agree(cuda):
    """The streaming and the row-per-lane SpMM kernels of the engine run the same
    recurrences: same iteration counts (within one) and the same solutions to rounding."""
    from tensors_amd import _capi as capi

    sysl = systems(6, seed=99)
    before = capi.lib().tg_tune_get(b"batch_stream")
    try:
        capi.call("tg_tune_set", b"batch_stream", 1)
        Xs, rs = run_batch(sysl, cuda, False, 1e-10)
        capi.call("tg_tune_set", b"batch_stream", 0)
        Xr, rr = run_batch(sysl, cuda, False, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_stream", before)
    for i in range(len(sysl)):
        assert abs(rs[i]["iterations"] - rr[i]["iterations"]) <= 1
        assert rel_err(Xs[i], Xr[i]) < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_split_and_fused_direction_forms_are_bitwise_equal(cuda, mixed):
    """The split iteration (direction pass, SpMM of p, update) and the fused one (direction
    rebuilt inside the SpMM) compute p, u and every partial with the same operations in the
    same order: identical iterates, iteration counts and replacements."""
    from tensors_amd import _capi as capi

    sysl = systems(6, seed=123)
    before = (capi.lib().tg_tune_get(b"batch_split"), capi.lib().tg_tune_get(b"batch_stream"))
    try:
        capi.call("tg_tune_set", b"batch_stream", 0)
        capi.call("tg_tune_set", b"batch_split", 1)
        Xs, rs = run_batch(sysl, cuda, mixed, 1e-10)
        capi.call("tg_tune_set", b"batch_split", 0)
        Xf, rf = run_batch(sysl, cuda, mixed, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_split", before[0])
        capi.call("tg_tune_set", b"batch_stream", before[1])
    for i in range(len(sysl)):
        assert rs[i]["iterations"] == rf[i]["iterations"]
        assert np.array_equal(Xs[i], Xf[i])
