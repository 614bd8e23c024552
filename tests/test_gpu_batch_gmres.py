"""Batched GMRES engine (tg_gbatch_*): every system of an uneven batch against the
oracle's GMRES at a fixed number of cycles (tol = 0, so every cycle runs its m steps) at
1e-12, converged solves against the oracle and the host residual, and a system's result
is bitwise the same alone, in a mix and when it joins while others run."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.batch import BatchGMRES

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["delayed", "cgs2"], autouse=True)
def gbatch_form(request):
    """Every test with both step forms of the engine (knob gbatch_delayed, read when an
    engine is created): the delayed second Gram-Schmidt pass and plain CGS2."""
    from tensors_amd import _capi as capi

    before = int(capi.lib().tg_tune_get(b"gbatch_delayed"))
    capi.call("tg_tune_set", b"gbatch_delayed", request.param)
    yield request.param
    capi.call("tg_tune_set", b"gbatch_delayed", before)


def systems(count, seed=5005, gmax=70, ms=(5, 12, 30)):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        g = int(rng.integers(6, gmax))
        pe = float(np.exp(rng.uniform(0, np.log(100))))
        rp, ci, val = gen.convection_diffusion_2d(g, pe)
        n = g * g
        k = int(rng.integers(1, 4))
        m = int(rng.choice(ms))
        B = gen.uniform(seed + 17 * s, n * k).reshape(n, k)
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if s % 3 == 1 else None
        out.append((rp.astype(np.int32), ci, val, d, B, n, k, m))
    return out


def to_dev(sysd, dev):
    rp, ci, val, d, B, n, k, m = sysd
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    return (T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64),
            None if d is None else T(d, torch.float64), T(B, torch.float64),
            torch.zeros(n, k, dtype=torch.float64, device=dev))


def run(sysl, dev, tol, max_restarts, order=None, staged=False):
    bt = BatchGMRES()
    devs = [to_dev(s, dev) for s in sysl]
    ids = []
    idx = list(order) if order is not None else list(range(len(sysl)))
    half = len(idx) // 2
    for jj, i in enumerate(idx):
        rp, ci, val, d, B, X = devs[i]
        ids.append((i, bt.add(rp, ci, val, B, X, tol, sysl[i][7], max_restarts, diag=d)))
        if staged and jj == half:
            bt.step(1)  # the rest join while the first half runs
    bt.run(1)
    res = {i: bt.result(sid) for i, sid in ids}
    X = {i: devs[i][5].cpu().numpy() for i, _ in ids}
    bt.close()
    return X, res


@pytest.mark.parametrize("cycles", [1, 2])
def test_batch_gmres_fixed_cycles_match_oracle(cuda, cycles):
    # restart lengths 5, 12 and 30, and two systems of 40 and 63 (the largest basis
    # bucket of the step kernels) in the same batch
    sysl = systems(9, gmax=40) + systems(2, seed=5105, gmax=40, ms=(40, 63))
    # and blocks of 5 and 8 right-hand sides (eight column lanes in the reductions)
    for kk, g in ((5, 17), (8, 23)):
        rp, ci, val = gen.convection_diffusion_2d(g, 7.0)
        n = g * g
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if kk == 8 else None
        sysl.append((rp.astype(np.int32), ci, val, d, gen.uniform(5200 + kk, n * kk).reshape(n, kk), n, kk, 12))
    X, res = run(sysl, cuda, 0.0, cycles - 1)
    for i, (rp, ci, val, d, B, n, k, m) in enumerate(sysl):
        Xo = np.zeros((n, k))
        ok, it, rs, rr = oracle.gmres(n, rp, ci, val, B, Xo, k, m, 0.0, cycles - 1, prec_diag=d)
        r = res[i]
        print(f"system {i}: n {n} k {k} m {m} prec {d is not None}", flush=True)
        assert r["finished"] and not r["succeeded"] and not ok
        assert r["iterations"] == it == min(m, n + 1) and r["restarts"] == rs == cycles - 1
        assert rel_err(X[i], Xo) <= 1e-12
        # residual estimates in units of ||P b|| (a residual of 1e-8 ||P b|| is known to
        # rounding in those units, not to 1e-12 of itself)
        assert np.max(np.abs(r["rel_res"] - rr)) <= 1e-12


def test_batch_gmres_converges_like_oracle(cuda):
    sysl = systems(12, ms=(20, 30))
    X, res = run(sysl, cuda, 1e-8, 200)
    for i, (rp, ci, val, d, B, n, k, m) in enumerate(sysl):
        r = res[i]
        Xo = np.zeros((n, k))
        ok, it, rs, rr = oracle.gmres(n, rp, ci, val, B, Xo, k, m, 1e-8, 200, prec_diag=d)
        assert r["succeeded"] and ok, (i, r)
        assert np.all(r["rel_res"] <= 1e-8)
        # the same restart count (iterations may differ by a step at the threshold)
        assert abs(r["restarts"] - rs) <= 1 and rel_err(X[i], Xo) <= 1e-6
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        R = B - A @ X[i]
        if d is not None:
            true = np.linalg.norm(d[:, None] * R, axis=0) / np.linalg.norm(d[:, None] * B, axis=0)
        else:
            true = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(true <= 1.5e-8), (i, true)


def test_batch_gmres_alone_equals_mix_and_mid_flight(cuda):
    sysl = systems(8, seed=6006, ms=(12, 30))
    mix, rmix = run(sysl, cuda, 1e-9, 100)
    staged, rst = run(sysl, cuda, 1e-9, 100, order=[7, 3, 5, 1, 0, 2, 4, 6], staged=True)
    for i in range(len(sysl)):
        alone, ra = run([sysl[i]], cuda, 1e-9, 100)
        assert np.array_equal(alone[0], mix[i]), i
        assert np.array_equal(staged[i], mix[i]), i
        assert ra[0]["iterations"] == rmix[i]["iterations"] and ra[0]["restarts"] == rmix[i]["restarts"]
        assert np.array_equal(ra[0]["rel_res"], rmix[i]["rel_res"])


def test_batch_gmres_special_cases(cuda):
    rp, ci, val = gen.convection_diffusion_2d(10, 5.0)
    n = 100
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dt)
    rpd, cid, vd = T(rp.astype(np.int32), torch.int32), T(ci, torch.int32), T(val, torch.float64)
    bt = BatchGMRES()
    Bz = torch.zeros(n, 2, dtype=torch.float64, device=cuda)
    Xz = torch.full((n, 2), 3.0, dtype=torch.float64, device=cuda)
    Bn = T(gen.uniform(1, 2 * n).reshape(n, 2), torch.float64)
    Bn[4, 1] = float("nan")
    Xn = torch.full((n, 2), 3.0, dtype=torch.float64, device=cuda)
    Bt = T(gen.uniform(2, n).reshape(n, 1), torch.float64)  # a 1 x 1 system: m = min(30, 2)
    one = (T(np.array([0, 1], np.int32), torch.int32), T(np.array([0], np.int32), torch.int32),
           T(np.array([4.0]), torch.float64))
    Xt = torch.zeros(1, 1, dtype=torch.float64, device=cuda)
    a = bt.add(rpd, cid, vd, Bz, Xz, 1e-8)
    b = bt.add(rpd, cid, vd, Bn, Xn, 1e-8)
    c = bt.add(one[0], one[1], one[2], Bt[:1], Xt, 1e-12)
    bt.run(1)
    ra, rb, rc = bt.result(a), bt.result(b), bt.result(c)
    assert ra["succeeded"] and torch.all(Xz == 0.0)  # zero right-hand side: X = 0
    assert not rb["succeeded"] and torch.all(Xn == 3.0)  # NaN: nothing written
    assert rc["succeeded"] and abs(Xt.item() - Bt[0, 0].item() / 4.0) <= 1e-15
    bt.close()


@pytest.fixture
def big_tiles_knob():
    from tensors_amd import _capi as capi

    before = int(capi.lib().tg_tune_get(b"gbatch_big_tiles"))
    yield lambda v: capi.call("tg_tune_set", b"gbatch_big_tiles", int(v))
    capi.call("tg_tune_set", b"gbatch_big_tiles", before)


def test_batch_gmres_big_system_totals_are_bitwise_neutral(cuda, big_tiles_knob):
    """Systems of more than gbatch_big_tiles tiles get their reduction totals from one
    workgroup per system (gb_totals_kernel) instead of every tile workgroup adding all
    tiles' partials: the same sums in the same order, so the results are bitwise those of
    the per-workgroup totals.  A mix with a 300^2 system (88 tiles) and small ones, run
    with the default threshold (the large one big), with every system big (0) and with
    none (2^30); the large system also against the host residual."""
    sysl = systems(5, seed=7007, ms=(12, 30))
    rp, ci, val = gen.convection_diffusion_2d(300, 20.0)
    n = 300 * 300
    sysl.append((rp.astype(np.int32), ci, val, None, gen.uniform(7100, n * 2).reshape(n, 2), n, 2, 30))
    out = {}
    for thr in (64, 0, 1 << 30):
        big_tiles_knob(thr)
        print(f"batched GMRES, big-system threshold {thr} tiles", flush=True)
        out[thr] = run(sysl, cuda, 1e-9, 200)
    for thr in (0, 1 << 30):
        for i in range(len(sysl)):
            assert np.array_equal(out[thr][0][i], out[64][0][i]), (thr, i)
            assert out[thr][1][i]["iterations"] == out[64][1][i]["iterations"]
            assert out[thr][1][i]["restarts"] == out[64][1][i]["restarts"]
            assert np.array_equal(out[thr][1][i]["rel_res"], out[64][1][i]["rel_res"])
    Xl, rl = out[64][0][len(sysl) - 1], out[64][1][len(sysl) - 1]
    A = sp.csr_matrix((val, ci, rp), shape=(n, n))
    B = sysl[-1][4]
    assert rl["succeeded"]
    assert np.max(np.linalg.norm(B - A @ Xl, axis=0) / np.linalg.norm(B, axis=0)) <= 1.5e-9
