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
    assert len(table) == len(autotune.candidates(4)) + 3 and all(t > 0 for _, t in table)
    best, _ = autotune.tune_spmm(A, X, Y, reps=2, keep=True)
    assert all(autotune.get_knobs()[k] == v for k, v in best.items())
    autotune.set_knobs(before)
