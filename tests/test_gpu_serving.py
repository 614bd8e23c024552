"""Serving loop: CG and GMRES requests with host arrays, admitted under a device-memory
budget.  Each result equals the same request solved alone (bitwise), meets its
tolerance against a host residual, GMRES agrees with the oracle GMRES, the budget is
never exceeded, and a request larger than the budget is refused with a status."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import rel_err
from tensors_amd import _capi as capi
from tensors_amd import generators as gen
from tensors_amd.serving import Server

pytestmark = pytest.mark.gpu


def requests(seed=5005, count=12):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        if s % 3 == 2:
            g = int(rng.integers(24, 64))
            pe = float(np.exp(rng.uniform(0, np.log(100))))
            rp, ci, val = gen.convection_diffusion_2d(g, pe)
            n = g * g
            B = gen.uniform(seed + s, n).reshape(n, 1)
            out.append(("gmres", rp, ci, val, B, None))
        else:
            dims = 2 if s % 2 else 3
            side = int(rng.integers(8, 40 if dims == 2 else 14))
            rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=seed + s)
            n = len(rp) - 1
            k = int(rng.integers(1, 9))
            B = gen.uniform(seed + 100 + s, n * k).reshape(n, k)
            d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
            out.append(("cg", rp, ci, val, B, d))
    return out


def submit(srv, req):
    kind, rp, ci, val, B, d = req
    if kind == "cg":
        return srv.submit_cg(rp, ci, val, B, tol=1e-10, max_iter=20000, inv_diag=d, mixed=True)
    return srv.submit_gmres(rp, ci, val, B, tol=1e-8, restart=30, max_restarts=100)


@pytest.fixture(params=[(1, 1 << 30, 0), (0, 1 << 30, 0), (1, 2000, 0), (0, 1 << 30, 1), (1, 2000, 1)],
                ids=["gmres_batched", "gmres_workers", "gmres_split", "gmres_workers_mixed", "gmres_split_mixed"])
def gmres_mode(request):
    """Serving GMRES requests through the batched GMRES engine (every request here
    fits it), on per-request worker streams, or split by size (systems of up to 2000
    rows batched, the larger ones on workers, as the default split does at 131072
    rows); the worker-stream requests with tg_gmres_solve or the mixed solver
    (serve_gmres_mixed); restores the knobs afterwards."""
    keys = (b"serve_gmres_batch", b"serve_gmres_batch_rows", b"serve_gmres_mixed")
    before = [int(capi.lib().tg_tune_get(k)) for k in keys]
    for k, v in zip(keys, request.param):
        capi.call("tg_tune_set", k, v)
    yield request.param
    for k, v in zip(keys, before):
        capi.call("tg_tune_set", k, v)


def solve_all(reqs, budget=8 << 30, workers=3):
    srv = Server(budget, workers)
    ids = [submit(srv, r) for r in reqs]
    res = srv.drain(ids, 300)
    stats = srv.stats()
    srv.close()
    return [res[i] for i in ids], stats


def test_serving_results_meet_tolerance_and_match_oracle(cuda, gmres_mode):
    reqs = requests()
    res, stats = solve_all(reqs)
    assert stats["bytes_in_use"] == 0 and stats["running"] == 0
    for (kind, rp, ci, val, B, d), r in zip(reqs, res):
        assert r["status"] == 0
        n = len(rp) - 1
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        X = r["X"]
        true = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
        if kind == "cg":
            assert r["succeeded"]
            assert np.all(true <= 1.5e-10), true
        else:
            Xo = np.zeros_like(B)
            ok, it, rst, rr = oracle.gmres(n, rp, ci, val, B, Xo, 1, 30, 1e-8, 100)
            assert bool(r["succeeded"]) == bool(ok)
            if ok:
                assert np.all(true <= 1.5e-8), true
                assert rel_err(X, Xo) < 1e-5


def test_serving_alone_equals_mix_and_budget_holds(cuda, gmres_mode):
    reqs = requests(seed=5105, count=9)
    mix, _ = solve_all(reqs)
    sizes = [Server.request_bytes(r[0], len(r[1]) - 1, len(r[2]), r[4].shape[1], r[5] is not None, True, 30)
             for r in reqs]
    budget = 2 * max(sizes)
    tight, stats = solve_all(reqs, budget=budget, workers=2)
    assert 0 < stats["peak_bytes"] <= budget
    for j, r in enumerate(reqs):
        alone, _ = solve_all([r])
        assert np.array_equal(alone[0]["X"], mix[j]["X"]), j
        assert np.array_equal(tight[j]["X"], mix[j]["X"]), j
        assert alone[0]["iterations"] == mix[j]["iterations"]


def test_serving_refuses_request_over_budget(cuda):
    kind, rp, ci, val, B, d = requests(count=1)[0]
    srv = Server(1 << 16, 1)
    sid = submit(srv, (kind, rp, ci, val, B, d))
    r = srv.drain([sid], 60)[sid]
    srv.close()
    assert r["status"] == capi.ERR_WORKSPACE


def test_serving_edge_requests(cuda, gmres_mode):
    """Tiny CG requests (n = 1 and 2, up to 8 right-hand sides), a tiny GMRES request (3 x 3
    grid), and CG / GMRES requests whose right-hand side holds a NaN, in one server: the tiny
    ones reach their tolerance, the NaN ones come back as not succeeded with status 0, and
    nothing stays admitted."""
    def tri(n):
        A = sp.diags([np.full(n - 1, -1.0), np.full(n, 2.2), np.full(n - 1, -1.0)], [-1, 0, 1], format="csr") \
            if n > 1 else sp.csr_matrix(np.array([[2.0]]))
        A.sort_indices()
        return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, 1.0 / A.diagonal()

    reqs = []
    for n, k in ((1, 1), (2, 8), (3, 3)):
        rp, ci, val, d = tri(n)
        reqs.append(("cg", rp, ci, val, gen.uniform(70 + n, n * k).reshape(n, k), d))
    rp, ci, val = gen.convection_diffusion_2d(3, 5.0)
    reqs.append(("gmres", rp, ci, val, gen.uniform(80, 9).reshape(9, 1), None))
    rp, ci, val, d = tri(50)
    Bn = gen.uniform(81, 100).reshape(50, 2)
    Bn[7, 1] = np.nan
    reqs.append(("cg", rp, ci, val, Bn, d))
    rp, ci, val = gen.convection_diffusion_2d(10, 5.0)
    Bg = gen.uniform(82, 100).reshape(100, 1)
    Bg[3, 0] = np.nan
    reqs.append(("gmres", rp, ci, val, Bg, None))
    res, stats = solve_all(reqs, workers=2)
    assert stats["bytes_in_use"] == 0 and stats["running"] == 0
    for j, ((kind, rp, ci, val, B, d), r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0, (j, r["status"])
        if j < 4:
            n = len(rp) - 1
            A = sp.csr_matrix((val, ci, rp), shape=(n, n))
            true = np.linalg.norm(B - A @ r["X"], axis=0) / np.linalg.norm(B, axis=0)
            assert r["succeeded"], (j, r)
            assert np.all(true <= (1.5e-10 if kind == "cg" else 1.5e-8)), (j, true)
        else:
            assert not r["succeeded"], (j, r)
