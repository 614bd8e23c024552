"""The solver workloads at their full sizes, through size-independent properties.

* 3D 7-point Laplacian on 128^3 (+ 1e-2 shift, n = 2 097 152) with 32 right-hand sides:
  FP64 CG and mixed CG (FP32 inner, FP64 residual replacement) both reach the 1e-10
  relative residual, checked as the true residual ||B - A X|| / ||B|| on the host, and
  the two solutions agree to 1e-10.
* The 20 largest systems of the uneven batch workload (2 000 SPD weighted grid
  Laplacians, n log-uniform in [1e3, 2e5], seed 4004; n up to 2e5) in one batch, each in
  all three modes: FP64 with the reference stop test, mixed, and FP64 with the
  true-residual stop; true residuals checked on the host.
* A Poisson-arrival serving run of CG requests (as above) and GMRES(30) requests on 2D
  upwind convection-diffusion with n in [1e4, 1e6]: every result meets its tolerance and
  equals, bit for bit, the same request served alone.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.batch import BatchCG
from tensors_amd.serving import Server
from tensors_amd.solvers import ConjugateGradient, cg_solve_mixed
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_laplacian_128_cubed_32_rhs_fp64_and_mixed(cuda):
    g, k = 128, 32
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    B = gen.uniform(3003, n * k).reshape(n, k)
    A = MatrixCSR(rp.astype(np.int32), ci, val, n, n)
    As = sp.csr_matrix((val, ci, rp), shape=(n, n))
    Bd = torch.from_numpy(B).to(cuda)
    bn = np.linalg.norm(B, axis=0)

    t = time.perf_counter()
    X64 = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    cg = ConjugateGradient(n, 2000, k)
    assert cg(A, None, 1.0, Bd, 0.0, X64, 1e-10)
    say(f"128^3 x 32: FP64 CG {cg.IterationCount()} iterations in {time.perf_counter() - t:.2f}s")
    Xm = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    t = time.perf_counter()
    ok, outer, inner, rr = cg_solve_mixed(A, Bd, Xm, tol=1e-10)
    say(f"128^3 x 32: mixed CG {inner} FP32 iterations, {outer} replacements in {time.perf_counter() - t:.2f}s")
    assert ok and np.all(rr <= 1e-10)
    X64h, Xmh = X64.cpu().numpy(), Xm.cpu().numpy()
    true64 = np.linalg.norm(B - As @ X64h, axis=0) / bn
    truem = np.linalg.norm(B - As @ Xmh, axis=0) / bn
    agree = rel_err(Xmh, X64h)
    say(f"true residuals: FP64 max {true64.max():.3e}, mixed max {truem.max():.3e}; mixed vs FP64 {agree:.3e}")
    # FP64 stops on the recursive residual (the reference's test), mixed on the true one
    assert np.all(true64 <= 1.5e-10) and np.all(truem <= 1.0000001e-10)
    assert np.allclose(truem, rr, rtol=1e-6, atol=0)
    assert agree <= 1e-10


def batch_specs(count=2000, seed=4004):
    """The uneven batch workload's systems (dims, side, nrhs), drawn as the benchmark draws them."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n_t = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        dims = 2 if rng.random() < 0.5 else 3
        side = max(2, int(round(n_t ** (1.0 / dims))))
        specs.append((dims, side, int(rng.integers(1, 9))))
    return specs


def test_batch_largest_systems_all_modes(cuda):
    specs = batch_specs()
    largest = sorted(range(len(specs)), key=lambda i: specs[i][1] ** specs[i][0], reverse=True)[:20]
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dt)
    bt = BatchCG()
    cases = []
    for i in largest:
        dims, side, k = specs[i]
        rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=9000 + i)
        n = len(rp) - 1
        As = sp.csr_matrix((val, ci, rp), shape=(n, n))
        d = 1.0 / As.diagonal()
        B = gen.uniform(40040 + i, n * k).reshape(n, k)
        dv = (T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64), T(d, torch.float64), T(B, torch.float64))
        for mode in (0, 1, 2):
            X = torch.zeros(n, k, dtype=torch.float64, device=cuda)
            sid = bt.add(dv[0], dv[1], dv[2], dv[4], X, 1e-10, 20000, diag=dv[3], mixed=mode == 1,
                         true_residual=mode == 2)
            cases.append((sid, mode, n, k, As, B, X, dv))
    assert max(c[2] for c in cases) >= 1.5e5
    say(f"batch: {len(cases)} systems (20 largest x 3 modes), n {min(c[2] for c in cases)}..{max(c[2] for c in cases)}")
    t = time.perf_counter()
    bt.run(16)
    say(f"batch: done in {time.perf_counter() - t:.2f}s")
    worst0 = 0.0
    for sid, mode, n, k, As, B, X, _ in cases:
        r = bt.result(sid)
        assert r["finished"] and r["succeeded"], (sid, mode, r)
        assert np.all(r["rel_res"] <= 1e-10), (sid, mode, r)
        true = np.linalg.norm(B - As @ X.cpu().numpy(), axis=0) / np.linalg.norm(B, axis=0)
        if mode == 0:
            # the reference's test is on the preconditioned recursive residual sqrt(r.Pr / b.Pb)
            worst0 = max(worst0, true.max())
            assert np.all(true <= 1e-7), (sid, true)
        else:
            assert np.all(true <= 1.0000001e-10), (sid, mode, true)
            assert np.allclose(true, r["rel_res"], rtol=1e-6, atol=0)
    say(f"batch: mode 0 worst true residual {worst0:.3e}")
    bt.close()


def serve_requests(count=12, seed=5005):
    """CG requests shaped like the batch systems and GMRES(30) requests on 2D upwind
    convection-diffusion, n log-uniform in [1e4, 1e6], Peclet log-uniform in [1, 100]."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        if s % 2 == 0:
            n_t = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
            dims = 2 if rng.random() < 0.5 else 3
            side = max(2, int(round(n_t ** (1.0 / dims))))
            k = int(rng.integers(1, 9))
            rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=seed * 10 + s)
            n = len(rp) - 1
            d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
            B = gen.uniform(seed * 100 + s, n * k).reshape(n, k)
            out.append(("cg", rp.astype(np.int32), ci, val, B, d))
        else:
            n_t = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
            g = max(2, int(round(np.sqrt(n_t))))
            pe = float(np.exp(rng.uniform(0.0, np.log(100.0))))
            rp, ci, val = gen.convection_diffusion_2d(g, pe)
            B = gen.uniform(seed * 100 + s, g * g).reshape(g * g, 1)
            out.append(("gmres", rp.astype(np.int32), ci, val, B, None))
    # one request at the top of the GMRES size range
    rp, ci, val = gen.convection_diffusion_2d(1000, 10.0)
    out.append(("gmres", rp.astype(np.int32), ci, val, gen.uniform(5999, 10 ** 6).reshape(10 ** 6, 1), None))
    return out


def submit(srv, r):
    kind, rp, ci, val, B, d = r
    if kind == "cg":
        return srv.submit_cg(rp, ci, val, B, tol=1e-10, max_iter=20000, inv_diag=d, mixed=True)
    return srv.submit_gmres(rp, ci, val, B, tol=1e-8, restart=30, max_restarts=100)


def test_serving_poisson_arrivals_alone_equals_mix(cuda):
    reqs = serve_requests()
    assert max(len(r[1]) - 1 for r in reqs if r[0] == "gmres") >= 10 ** 6
    assert min(len(r[1]) - 1 for r in reqs if r[0] == "gmres") >= 10 ** 4
    rng = np.random.default_rng(5006)
    gaps = rng.exponential(1.0 / 20.0, len(reqs))  # 20 requests per second offered
    srv = Server(200 << 30, 3)
    ids, t0 = [], time.perf_counter()
    for r, a in zip(reqs, np.cumsum(gaps)):
        while time.perf_counter() - t0 < a:
            time.sleep(0.001)
        ids.append(submit(srv, r))
    mix = srv.drain(ids, 600)
    stats = srv.stats()
    srv.close()
    say(f"serving: {len(ids)} Poisson arrivals served in {time.perf_counter() - t0:.2f}s, "
        f"peak admitted {stats['peak_bytes'] / 1e9:.2f} GB")
    assert stats["bytes_in_use"] == 0 and stats["running"] == 0
    for j, (r, sid) in enumerate(zip(reqs, ids)):
        kind, rp, ci, val, B, d = r
        res = mix[sid]
        assert res["status"] == 0 and res["succeeded"], (j, kind, res)
        n = len(rp) - 1
        As = sp.csr_matrix((val, ci, rp), shape=(n, n))
        true = np.linalg.norm(B - As @ res["X"], axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(true <= (1.0000001e-10 if kind == "cg" else 1.5e-8)), (j, kind, true)
        alone_srv = Server(200 << 30, 1)
        aid = submit(alone_srv, r)
        alone = alone_srv.drain([aid], 600)[aid]
        alone_srv.close()
        assert np.array_equal(alone["X"], res["X"]), (j, kind)
        assert alone["iterations"] == res["iterations"] and alone["restarts"] == res["restarts"], (j, kind)
    say("serving: every request equals its alone run bit for bit")
