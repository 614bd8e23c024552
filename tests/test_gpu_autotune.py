"""Every SpMM shape the autotuner can pick computes the same product: for each knob
setting (nrhs = 1 kernels and team widths, nrhs >= 2 lane-group widths) the result
matches the FP64 oracle within 1e-12 relative, and tune_spmm leaves the fastest
setting in place and restores on request."""
import numpy as np
import pytest
import torch

import oracle
from helpers import random_csr_np, rel_err
from tensors_amd import autotune
from tensors_amd.sparse import MatrixCSR, spmm_raw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("density", [0.01, 0.002])
@pytest.mark.parametrize("nrhs", [1, 3, 8, 33])
def test_every_shape_matches_oracle(cuda, nrhs, density):
    m, n = 3000, 2500
    rp, ci, val = random_csr_np(m, n, density, seed=77 + nrhs, empty_rows=0.05)
    rng = np.random.default_rng(nrhs)
    X = rng.uniform(-1, 1, (n, nrhs))
    Yo = np.zeros((m, nrhs))
    oracle.spmm(rp, ci, val, m, n, 1.0, X, nrhs, 0.0, Yo, nrhs, nrhs)
    A = MatrixCSR(rp.astype(np.int32), ci, val, m, n)
    rpd, cid, vd = A.device_arrays(torch.float64)
    Xd = torch.from_numpy(X).to(cuda)
    before = autotune.get_knobs()
    try:
        for cfg in autotune.candidates(nrhs):
            autotune.set_knobs(cfg)
            Yd = torch.full((m, nrhs), float("nan"), dtype=torch.float64, device=cuda)
            spmm_raw(m, n, A.NonzeroCount(), rpd, cid, vd, 1.0, Xd, nrhs, 0.0, Yd, nrhs, nrhs)
            assert rel_err(Yd.cpu().numpy(), Yo) <= 1e-12, cfg
    finally:
        autotune.set_knobs(before)


def test_tune_spmm_keeps_best_and_restores(cuda):
    m = n = 20000
    rp, ci, val = random_csr_np(m, n, 0.002, seed=5)
    A = MatrixCSR(rp.astype(np.int32), ci, val, m, n)
    X = torch.rand(n, 4, dtype=torch.float64, device=cuda)
    Y = torch.empty(m, 4, dtype=torch.float64, device=cuda)
    before = autotune.get_knobs()
    best, table = autotune.tune_spmm(A, X, Y, reps=2, keep=False)
    assert autotune.get_knobs() == before
    assert len(table) == len(autotune.candidates(4)) + 4 and all(t > 0 for _, t in table)
    best, _ = autotune.tune_spmm(A, X, Y, reps=2, keep=True)
    assert all(autotune.get_knobs()[k] == v for k, v in best.items())
    autotune.set_knobs(before)


def test_tune_batch_keeps_best_and_every_form_matches(cuda):
    """tune_batch times every (tile rows, SpMM form) candidate, keeps the fastest and
    restores on request; the solution after a fixed number of FP64 iterations is bitwise
    the same for every SpMM form and within rounding across tile heights."""
    import scipy.sparse as sp

    from tensors_amd import generators as gen
    from tensors_amd.batch import BatchCG

    sysl = []
    for s, (dims, side) in enumerate(((2, 90), (3, 20), (2, 40))):
        rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=300 + s)
        n = len(rp) - 1
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
        B = gen.uniform(400 + s, n * 3).reshape(n, 3)
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dt)
        sysl.append((T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64), T(d, torch.float64),
                     T(B, torch.float64)))
    before = autotune.get_batch_knobs()
    cands = autotune.batch_candidates(tile_rows=(256, 1024), spmm_forms=(2, 0))
    best, table = autotune.tune_batch(sysl, iters=4, warm=1, candidates=cands, keep=False)
    assert autotune.get_batch_knobs() == before and len(table) == 4 and all(t > 0 for _, t in table)
    best, _ = autotune.tune_batch(sysl, iters=4, warm=1, candidates=cands, keep=True)
    assert autotune.get_batch_knobs() == best
    results = {}
    try:
        for cfg in cands:
            autotune.set_knobs(cfg)
            bt = BatchCG()
            Xs = [torch.zeros_like(s[4]) for s in sysl]
            for (rp, ci, val, d, B), X in zip(sysl, Xs):
                bt.add(rp, ci, val, B, X, 0.0, 30, diag=d)
            bt.run(8)
            bt.close()
            results[(cfg["batch_tile_rows"], cfg["batch_spmm"])] = [X.cpu().numpy() for X in Xs]
    finally:
        autotune.set_knobs(before)
    for t in (256, 1024):
        for a, b in zip(results[(t, 2)], results[(t, 0)]):
            assert np.array_equal(a, b)
    for a, b in zip(results[(256, 2)], results[(1024, 2)]):
        assert rel_err(a, b) <= 1e-12
